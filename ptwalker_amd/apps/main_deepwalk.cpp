// Drop-in for the reference's main() in src/core/deepwalk.cpp:44-63.
#include "../../include/ptw.h"
#include "cli.hpp"

int main(int argc, char **argv)
{
    return ptw_cli::run(PTW_APP_DEEPWALK, ptw_cli::Parser::WALK, argv[0], argc, argv);
}
