// Drop-in for the reference's main() in src/core/pagerank.cpp:93-112.
#include "../../include/ptw.h"
#include "cli.hpp"

int main(int argc, char **argv)
{
    return ptw_cli::run(PTW_APP_PAGERANK, ptw_cli::Parser::RWPT, argv[0], argc, argv);
}
