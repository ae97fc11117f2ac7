// The fbls executable: `fbls search ...` and `fbls batch ...` (fbls_cli).
#include "../../include/fbls.h"

int main(int argc, char** argv) { return fbls_cli(argc - 1, (const char* const*)(argv + 1)); }
