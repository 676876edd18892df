// main.cpp - the `nrg` executable: a thin wrapper over libnrg_engine.so.
extern "C" int nrg_main(int argc, char **argv);

int main(int argc, char **argv) { return nrg_main(argc, argv); }
