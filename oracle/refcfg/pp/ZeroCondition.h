/* Initial-condition selection for the pseudopotential reference build: the
 * reference picks its initialiser by editing Ark/Constant/ZeroCondition.h,
 * which main() includes textually (Ark/Cpp/main.cpp:130); this one selects
 * SCMP_Drop (Ark/Cpp/Preprocess.cpp:337).  Found first on the -I path. */
SCMP_Drop();
