// ks: k-hop sampling (LE/src/apps/khop/main.cu) on the MI355X.
#include "app_main.hpp"

int main(int argc, char* argv[]) { return grass_app::appMain<grass_app::KHopApp>(argc, argv); }
