// Drives the C++ facade the way KnightKing's apps and examples do: load a text
// edge list, run node2vec / deepwalk (two epochs) / ppr (start list) / a
// std::function-only config (must be refused) / a WalkProgram (the dynamic
// walk example as device code), dump PathSets to files; node2vec's paths are
// also taken in HBM by a device consumer and written out after a copy.
// Then metapath() on a binary Edge<WeightedMetaData> file with schemes written
// and read back through the scheme-file functions, the vertex type masks
// computed as KnightKing's metapath() computes them (process_vertices over
// csr into an alloc_vertex_array), a CSR sorted on the host through csr and
// walked, and get_uniform_dist_func() against the starts the device draws.
// usage: facade_karate <karate.txt> <out_prefix> <program.cpp>
#include <stdio.h>

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <fstream>
#include <sstream>
#include <string>

#include "knightking/deepwalk.hpp"
#include "knightking/metapath.hpp"
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
    {
        // karate as binary Edge<WeightedMetaData> records, type = (src + 2 dst) % 3
        std::vector<Edge<WeightedMetaData> > edges;
        FILE *t = fopen(argv[1], "r");
        if (!t) return 10;
        unsigned a, b;
        float x;
        while (fscanf(t, "%u %u %f", &a, &b, &x) == 3) {
            WeightedMetaData d;
            d.weight = x;
            d.meta_info = (int)((a + 2 * b) % 3);
            edges.push_back(Edge<WeightedMetaData>(a, b, d));
        }
        fclose(t);
        std::string bin = out + "_meta.data";
        FILE *bf = fopen(bin.c_str(), "wb");
        if (!bf || fwrite(edges.data(), sizeof(edges[0]), edges.size(), bf) != edges.size()) return 11;
        fclose(bf);
        WalkEngine<WeightedMetaData, MetapathState> graph;
        graph.load_graph(34, bin.c_str());
        if (graph.get_edge_num() != edges.size() || graph.get_local_vertex_begin() != 0 ||
            graph.get_local_vertex_end() != 34)
            return 12;
        std::vector<std::vector<std::vector<bool> > > schemes = {
            {{true, false, true}, {false, true, true}},
            {{true, true, false}},
            {{false, false, true}, {true, false, false}, {false, true, false}}};
        std::string sp = out + "_schemes.txt";
        write_metapath_schemes(schemes, sp.c_str());
        auto rs = read_metapath_schemes(sp.c_str());
        if (rs != schemes) return 13;
        // KnightKing src/apps/metapath.hpp:77-88: type masks through csr
        scheme_mask_t *vm = graph.alloc_vertex_array<scheme_mask_t>();
        graph.process_vertices<vertex_id_t>([&](vertex_id_t v) {
            vm[v] = 0;
            for (auto p = graph.csr->adj_lists[v].begin; p < graph.csr->adj_lists[v].end; p++)
                vm[v] |= (scheme_mask_t)(1 << p->data.get_meta());
            return 0;
        });
        FILE *mf = fopen((out + "_meta_vmask.txt").c_str(), "w");
        if (!mf) return 14;
        for (vertex_id_t v = 0; v < 34; v++)
            fprintf(mf, "%u %u %u\n", (unsigned)vm[v], graph.vertex_out_degree[v], graph.vertex_in_degree[v]);
        fclose(mf);
        graph.dealloc_vertex_array(vm);
        WalkConfig wc;
        wc.set_output_file((out + "_meta").c_str(), true);
        metapath(&graph, rs, 68, 12, &wc);   // engine seed 1
    }
    {
        // a CSR edited on the host as KnightKing's node2vec() does (std::sort of
        // every list through process_vertices, src/apps/node2vec.hpp:122-125),
        // then walked: the device walks the sorted lists
        WalkEngine<real_t, EmptyData> graph;
        graph.load_graph(34, argv[1], false, GF_Edgelist);
        vertex_id_t degsum = graph.process_vertices<vertex_id_t>([&](vertex_id_t v) {
            std::sort(graph.csr->adj_lists[v].begin, graph.csr->adj_lists[v].end,
                      [](const AdjUnit<real_t> x, const AdjUnit<real_t> y) { return x.neighbour < y.neighbour; });
            return graph.vertex_out_degree[v];
        });
        if (degsum != graph.get_edge_num()) return 15;
        WalkConfig wc;
        wc.set_output_file((out + "_sorted_dw").c_str(), true);
        deepwalk(&graph, 34, 12, &wc);   // seed 1
        // get_uniform_dist_func: the starts the device draws for the next call
        auto ud = graph.get_uniform_dist_func();
        std::vector<vertex_id_t> want(40);
        for (walker_id_t i = 0; i < 40; i++) want[i] = ud(i);
        std::ifstream f(argv[3]);
        std::stringstream ss;
        ss << f.rdbuf();
        WalkProgram prog(ss.str());
        prog.param[0] = 5;
        prog.max_length = 5;
        WalkerConfig<real_t, EmptyData> wconf(40, nullptr, nullptr, ud);
        std::vector<vertex_id_t> got;
        WalkConfig wc2;
        wc2.set_output_consumer([&](PathSet *ps) {
            for (walker_id_t i = 0; i < ps->path_num[0]; i++) got.push_back(ps->path_begin[0][i][0]);
            delete ps;
        });
        graph.random_walk(&wconf, &prog, &wc2);   // seed 2
        if (got != want) return 16;
        int distinct = 0;
        for (vertex_id_t v = 0; v < 34; v++) distinct += std::count(want.begin(), want.end(), v) > 0;
        if (distinct < 10) return 17;   // drawn, not walker id modulo |V| alone
    }
    {
        // epochs streamed to a device consumer: each epoch's rows copied on the
        // stream the consumer is handed, while the next epoch walks
        WalkEngine<real_t, EmptyData> graph;
        graph.load_graph(34, argv[1], false, GF_Edgelist);
        WalkConfig wc;
        wc.set_walk_rate(0.5);
        wc.set_output_file((out + "_stream_dw").c_str(), true);
        std::vector<std::vector<uint32_t> > rows_e, lens_e;
        std::vector<uint32_t> strides;
        wc.set_device_stream_consumer([&](const kkw_paths &p, void *stream) {
            rows_e.emplace_back((size_t)p.walker_num * p.stride);
            lens_e.emplace_back(p.walker_num);
            strides.push_back(p.stride);
            if (hipMemcpyAsync(rows_e.back().data(), p.vertices, rows_e.back().size() * 4, hipMemcpyDeviceToHost,
                               (hipStream_t)stream) != hipSuccess ||
                hipMemcpyAsync(lens_e.back().data(), p.lengths, lens_e.back().size() * 4, hipMemcpyDeviceToHost,
                               (hipStream_t)stream) != hipSuccess)
                throw std::runtime_error("copy failed");
        });
        deepwalk(&graph, 34, 12, &wc);   // seed 1, two epochs of 17
        if (rows_e.size() != 2) return 18;
        FILE *f = fopen((out + "_stream_dev.0").c_str(), "w");
        if (!f) return 19;
        for (size_t e = 0; e < rows_e.size(); e++)
            for (size_t i = 0; i < lens_e[e].size(); i++) {
                for (uint32_t k = 0; k < (lens_e[e][i] & 0x7fffffffu); k++)
                    fprintf(f, k ? " %u" : "%u", rows_e[e][i * strides[e] + k]);
                fprintf(f, "\n");
            }
        fclose(f);
    }
    printf("facade ok\n");
    return 0;
}
