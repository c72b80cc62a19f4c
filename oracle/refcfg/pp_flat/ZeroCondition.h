/* Initial-condition selection for a pseudopotential reference build that starts
 * from the flat liquid band: the reference picks its initialiser by editing
 * Ark/Constant/ZeroCondition.h, which main() includes textually
 * (Ark/Cpp/main.cpp:130); this one selects SCMP_FlatInterface
 * (Ark/Cpp/Preprocess.cpp:291-336).  Found first on the -I path.  Test
 * infrastructure only. */
SCMP_FlatInterface();
