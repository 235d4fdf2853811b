// The host-only parts of the C++ facade under -fsanitize=address,undefined:
// scheme files written and read back, scheme masks, a PathSet built and
// dumped, an EdgeContainer, the Timer and the host generator.  Needs no GPU.
#include <stdio.h>

#include <string>

#include "knightking/metapath.hpp"

int main(int argc, char **argv)
{
    if (argc != 2) return 2;
    std::string dir = argv[1];
    std::vector<std::vector<std::vector<bool> > > schemes = {
        {{true, false, true, false}, {false, true, true, true}}, {{true, true, false, false}}};
    write_metapath_schemes(schemes, (dir + "/schemes.txt").c_str());
    auto back = read_metapath_schemes((dir + "/schemes.txt").c_str());
    if (back != schemes) return 3;
    auto masks = get_scheme_mask(back);
    if (masks.size() != 2 || masks[0][0] != 5 || masks[0][1] != 14 || masks[1][0] != 3) return 4;
    PathSet ps;
    ps.seg_num = 1;
    ps.path_set = new vertex_id_t *[1];
    ps.walker_id = new walker_id_t *[1];
    ps.path_begin = new vertex_id_t **[1];
    ps.path_end = new vertex_id_t **[1];
    ps.path_length = new step_t *[1];
    ps.path_num = new walker_id_t[1];
    ps.path_set[0] = new vertex_id_t[5]{1, 2, 3, 4, 5};
    ps.walker_id[0] = new walker_id_t[2]{0, 1};
    ps.path_begin[0] = new vertex_id_t *[2]{ps.path_set[0], ps.path_set[0] + 3};
    ps.path_end[0] = new vertex_id_t *[2]{ps.path_set[0] + 3, ps.path_set[0] + 5};
    ps.path_length[0] = new step_t[2]{3, 2};
    ps.path_num[0] = 2;
    ps.dump((dir + "/paths.txt").c_str(), "w", true);
    EdgeContainer<real_t> ec;
    ec.adj_units = new AdjUnit<real_t>[4];
    ec.adj_lists = new AdjList<real_t>[2];
    ec.adj_lists[0].begin = ec.adj_units;
    ec.adj_lists[0].end = ec.adj_units + 1;
    StdRandNumGenerator gen(7);
    for (int i = 0; i < 1000; i++)
        if (gen.gen(10) >= 10 || gen.gen_float(2.0f) >= 2.0f) return 5;
    Timer t;
    if (t.duration() < 0) return 6;
    printf("sanitized facade run: ok\n");
    return 0;
}
