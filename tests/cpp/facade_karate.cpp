// Drives the C++ facade the way KnightKing's apps and examples do: load a text
// edge list, run node2vec / deepwalk (two epochs) / ppr (start list) / a
// std::function-only config (must be refused) / a WalkProgram (the dynamic
// walk example as device code), dump PathSets to files; node2vec's paths are
// also taken in HBM by a device consumer and written out after a copy.
// usage: facade_karate <karate.txt> <out_prefix> <program.cpp>
#include <stdio.h>

#include <hip/hip_runtime_api.h>

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
        int dev_calls = 0;
        uint32_t dev_stride = 0;
        std::vector<uint32_t> rows, lens;
        wc.set_device_output_consumer([&](const kkw_paths &p) {
            dev_calls++;
            if (!p.on_device) return;
            dev_stride = p.stride;
            rows.resize((size_t)p.walker_num * p.stride);
            lens.resize(p.walker_num);
            if (hipMemcpy(rows.data(), p.vertices, rows.size() * 4, hipMemcpyDeviceToHost) != hipSuccess ||
                hipMemcpy(lens.data(), p.lengths, lens.size() * 4, hipMemcpyDeviceToHost) != hipSuccess)
                rows.clear();
        });
        node2vec(&graph, conf, &wc);   // engine seed 1, first call
        if (dev_calls != 1 || lens.size() != 34 || rows.empty()) return 5;
        FILE *f = fopen((out + "_n2v_dev.0").c_str(), "w");
        if (!f) return 6;
        for (size_t i = 0; i < lens.size(); i++) {
            for (uint32_t k = 0; k < (lens[i] & 0x7fffffffu); k++)
                fprintf(f, k ? " %u" : "%u", rows[i * dev_stride + k]);
            fprintf(f, "\n");
        }
        fclose(f);
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
