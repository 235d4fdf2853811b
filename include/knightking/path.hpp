// path.hpp -- PathSet, the output a WalkConfig consumer receives
// (KnightKing include/path.hpp:46-109): seg_num segments of paths, each with
// walker ids, begin/end pointers and lengths (vertices per path).  This build
// returns one segment in walker-id order.
#pragma once

#include <assert.h>
#include <stdio.h>
#include <string.h>

#include "type.hpp"

struct PathSet {
    int seg_num;
    vertex_id_t **path_set;
    walker_id_t **walker_id;
    vertex_id_t ***path_begin;
    vertex_id_t ***path_end;
    step_t **path_length;
    walker_id_t *path_num;

    PathSet() : seg_num(0), path_set(nullptr), walker_id(nullptr), path_begin(nullptr), path_end(nullptr),
                path_length(nullptr), path_num(nullptr) {}
    PathSet(const PathSet &) = delete;
    PathSet &operator=(const PathSet &) = delete;
    ~PathSet()
    {
        for (int s = 0; s < seg_num; s++) {
            delete[] path_set[s];
            delete[] walker_id[s];
            delete[] path_begin[s];
            delete[] path_end[s];
            delete[] path_length[s];
        }
        delete[] path_set;
        delete[] walker_id;
        delete[] path_begin;
        delete[] path_end;
        delete[] path_length;
        delete[] path_num;
    }
    // include/path.hpp:85-109: optional "id len", then " v" per vertex.
    void dump(const char *output_path, const char *fopen_mode, bool with_head_info)
    {
        FILE *f = fopen(output_path, fopen_mode);
        assert(f != NULL);
        for (int s = 0; s < seg_num; s++)
            for (walker_id_t i = 0; i < path_num[s]; i++) {
                if (with_head_info) fprintf(f, "%u %u", walker_id[s][i], path_length[s][i]);
                for (step_t k = 0; k < path_length[s][i]; k++) fprintf(f, " %u", path_begin[s][i][k]);
                fprintf(f, "\n");
            }
        fclose(f);
    }
};
