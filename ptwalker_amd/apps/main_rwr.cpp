// Drop-in for the reference's main() in src/core/RWR.cpp:109-128.
#include "../../include/ptw.h"
#include "cli.hpp"

int main(int argc, char **argv)
{
    return ptw_cli::run(PTW_APP_RWR, ptw_cli::Parser::WALK, argv[0], argc, argv);
}
