// Drop-in for the reference's main() in src/core/ppr.cpp:121-139.
#include "../../include/ptw.h"
#include "cli.hpp"

int main(int argc, char **argv)
{
    return ptw_cli::run(PTW_APP_PPR, ptw_cli::Parser::RWPT, argv[0], argc, argv);
}
