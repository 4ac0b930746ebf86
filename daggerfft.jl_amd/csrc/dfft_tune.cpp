// dfft_tune.cpp -- schedule cache of the autotuner (the FFTW wisdom analogue,
// src/fft.jl:113-124,162).
#include <mutex>
#include <string>
#include <unordered_map>

#include "dfft_plan.hpp"

namespace dfft {

static std::mutex g_mu;
static std::unordered_map<std::string, Tuning> g_cache;

std::string tuning_key(const PassPlan &pp) {
    const AxisPass &a = pp.ap;
    return std::to_string(pp.prec) + "/" + std::to_string(a.op) + "/" + std::to_string(a.sp.L) + "/" +
           std::to_string(a.contig);
}

bool tuning_lookup(const PassPlan &pp, Tuning &t) {
    std::lock_guard<std::mutex> g(g_mu);
    auto it = g_cache.find(tuning_key(pp));
    if (it == g_cache.end()) return false;
    t = it->second;
    return true;
}

}  // namespace dfft
