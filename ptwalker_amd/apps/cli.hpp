// Command-line front end shared by the pagerank, PPR, RWR and deepwalk
// binaries.  The flag surface is the reference's (include/option.hpp):
//   -f text|binary  -g PATH            GraphOptionHelper        :86-109,206-225
//   -t N  -s N  --socket-mapping=L  --mem N
//                                      NumaOptionHelper         :132-204
//   -e N | -w N  --sp P  -l N  --ws N  --db N
//                                      RWPTWalkOptionHelper     :227-302
//   -e N | -w N  -l N (required)       WalkOptionHelper         :304-355
//   -p P  -q Q (required, node2vec)    Node2vecOptionHelper     :447-482
//   -h / --help                        OptionParser             :21-59
// plus opt-in flags of this engine: --seed, --source, --device, --out,
// --json, --semantics, --dangling, --undirected, --out-ids, and for PPR and
// RWR the batch mode --queries / --topk.
#pragma once

namespace ptw_cli {

enum class Parser { RWPT, WALK, NODE2VEC };  // NODE2VEC: WALK plus -p and -q (include/option.hpp:447-482)

// Runs one app end to end; returns the process exit status.
int run(int app, Parser parser, const char *prog, int argc, char **argv);

}  // namespace ptw_cli
