// fastnsg: command-line build tool (same flags and files as the reference's
// src/fastnsg.cpp); the build runs on the GPU.
#include "cli_common.h"

int main(int argc, char **argv) { return fkcli::run_cli(argc, argv, fkcli::NSG); }
