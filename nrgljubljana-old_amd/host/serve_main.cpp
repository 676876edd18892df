// serve_main.cpp - the `nrg-serve` executable: a thin wrapper over
// libnrg_engine.so (serving loop in host/serve.cpp).
extern "C" int nrg_serve_main(int argc, char **argv);

int main(int argc, char **argv) { return nrg_serve_main(argc, argv); }
