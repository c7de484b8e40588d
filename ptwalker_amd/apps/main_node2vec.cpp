// Drop-in for the reference's main() in src/core/node2vec.cpp:8-24 (flags of
// Node2vecOptionParser, include/option.hpp:471-482).
#include "../../include/ptw.h"
#include "cli.hpp"

int main(int argc, char **argv)
{
    return ptw_cli::run(PTW_APP_NODE2VEC, ptw_cli::Parser::NODE2VEC, argv[0], argc, argv);
}
