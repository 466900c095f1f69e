// homo3d command line (main.cpp:22-48 of the reference): flags as in cmdline.cpp:4-27, e.g.
//   homo3d -reso=64 -obj=bulk -init=randc -sym=reflect6 -vol=0.3 -N=300 -prefix=out/
extern "C" int homo3d_framework_main(int argc, char** argv);

int main(int argc, char** argv) { return homo3d_framework_main(argc, argv); }
