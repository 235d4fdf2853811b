// util.hpp -- the host helpers KnightKing's apps use around a walk
// (include/util.hpp:36-91): a wall-clock Timer and the per-thread random
// number generator interface that user closures call.  On this engine the
// walks draw from the device's Philox stream; this generator only serves host
// code (closures evaluated on the host, start distributions).
#pragma once

#include <chrono>
#include <random>

#include "type.hpp"

class RandNumGenerator {
public:
    virtual vertex_id_t gen(vertex_id_t upper_bound) = 0;   // uniform in [0, upper_bound)
    virtual float gen_float(float upper_bound) = 0;         // uniform in [0, upper_bound)
    virtual ~RandNumGenerator() {}
};

class StdRandNumGenerator : public RandNumGenerator {
    std::mt19937 mt_;

public:
    StdRandNumGenerator() : mt_(std::random_device()()) {}
    explicit StdRandNumGenerator(uint32_t seed) : mt_(seed) {}
    vertex_id_t gen(vertex_id_t upper_bound)
    {
        return std::uniform_int_distribution<vertex_id_t>(0, upper_bound - 1)(mt_);
    }
    float gen_float(float upper_bound) { return std::uniform_real_distribution<float>(0.0f, upper_bound)(mt_); }
};

class Timer {
    std::chrono::time_point<std::chrono::system_clock> start_ = std::chrono::system_clock::now();

public:
    void restart() { start_ = std::chrono::system_clock::now(); }
    double duration()
    {
        return std::chrono::duration<double>(std::chrono::system_clock::now() - start_).count();
    }
    static double current_time()
    {
        return std::chrono::duration<double>(std::chrono::system_clock::now().time_since_epoch()).count();
    }
};
