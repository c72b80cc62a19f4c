/* Alternative flag set for building the reference's pseudopotential model
 * (Ark/Lib/D2Q9PP.cpp) with the UW face scheme: the flags of refcfg/pp/flags.h
 * with _FLUX_SCHEME_CD_ARK replaced by _FLUX_SCHEME_UW_ARK
 * (Ark/Constant/ZeroFlip.h:27-35).  Force-included ahead of ZeroFlip.h, whose
 * include guard it defines.  Test infrastructure only. */
#ifndef _ZERO_FLIP_H
#define _ZERO_FLIP_H
#include "ZeroMacros.h"
#define _CARTESIAN_MESH_FLIP
#define _MESHTYPE_ARK "Car"
#define _FLUX_SCHEME_UW_ARK "UW"
#define _FLUX_SCHEME_ARK _FLUX_SCHEME_UW_ARK
#define _BC_ARK "Periodic"
#define _EOS_MODEL_ARK "CarnahanStarling"
#define _MESHFILE_NAME_ARK "128_128_Car_Periodic_Square_LBM"
#define _PERIODIC_12_8_BCs_FLIP 12
#define _ARK_PSEUDOPSI_FLIP
#define _ARK_FORCE_FLIP
#define _ARK_FORCE_MODEL "He-Shan-Doolean"
#define _ARK_MOMENTUM_FLIP
#define _ARK_ISOTHERMAL_FLIP
#endif
