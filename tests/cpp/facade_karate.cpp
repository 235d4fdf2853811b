// Drives the C++ facade the way KnightKing's apps and examples do: load a text
// edge list, run node2vec / deepwalk (two epochs) / ppr (start list) / a
// std::function-only config (must be refused) / a WalkProgram (the dynamic
// walk example as device code), dump PathSets to files.
// usage: facade_karate <karate.txt> <out_prefix> <program.cpp>
#include <stdio.h>

#include <fstream>
#include <sstream>
#include <string>

#include "knightking/deepwalk.hpp"
#include "knightking/node2vec.hpp"
#include "knightking/ppr.hpp"

int main(int argc, char **argv)
{
    if (argc != 4) return 2;
    std::string out = argv[2];
    {
        WalkEngine<real_t, Node2vecState> graph;
        graph.load_graph(34, argv[1], false, GF_Edgelist);
        Node2vecConf conf;
        conf.p = 2;
        conf.q = 0.5;
        conf.walker_num = 34;
        conf.walk_length = 10;
        WalkConfig wc;
        wc.set_output_file((out + "_n2v").c_str());
        node2vec(&graph, conf, &wc);   // engine seed 1, first call
    }
    {
        WalkEngine<real_t, EmptyData> graph;
        graph.load_graph(34, argv[1], false, GF_Edgelist);
        WalkConfig wc;
        wc.set_output_file((out + "_dw").c_str(), true);
        wc.set_walk_rate(0.5);
        deepwalk(&graph, 34, 12, &wc);   // seed 1
        std::vector<vertex_id_t> starts = {0, 33};
        WalkConfig wc2;
        size_t got = 0;
        wc2.set_output_consumer([&](PathSet *ps) {
            got += ps->path_num[0];
            ps->dump((out + "_ppr.0").c_str(), "w", true);
            delete ps;
        });
        ppr(&graph, 20, 0.15f, &starts, &wc2);   // seed 2 (second call on this engine)
        if (got != 20) return 3;
        // a config built from callbacks alone has no device form
        WalkerConfig<real_t, EmptyData> wconf(5);
        TransitionConfig<real_t, EmptyData> tconf([](Walker<EmptyData> &, vertex_id_t) { return 0.875f; });
        try {
            graph.random_walk(&wconf, &tconf);
            return 4;
        } catch (const std::logic_error &) {
        }
    }
    {
        // src/examples/dynamic_walk.cpp as a WalkProgram, two epochs, with head info
        WalkEngine<real_t, EmptyData> graph;
        graph.load_graph(34, argv[1], false, GF_Edgelist);
        std::ifstream f(argv[3]);
        std::stringstream ss;
        ss << f.rdbuf();
        WalkProgram prog(ss.str());
        prog.param[0] = 10;     // walk length
        prog.max_length = 10;
        prog.use_weights = true;
        WalkerConfig<real_t, EmptyData> wconf(68);
        WalkConfig wc;
        wc.set_output_file((out + "_prog").c_str(), true);
        wc.set_walk_rate(0.5);
        graph.random_walk(&wconf, &prog, &wc);   // seed 1
    }
    printf("facade ok\n");
    return 0;
}
