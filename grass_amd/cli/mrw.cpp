// mrw: multi-dimensional random walks (LE/src/apps/multiRW/main.cu) on the MI355X.
#include "app_main.hpp"

int main(int argc, char* argv[]) { return grass_app::appMain<grass_app::MultiRWApp>(argc, argv); }
