// Host-side geometry plan of the estimator (internal to libbarbie_amd).
#pragma once
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "barbie.h"

namespace bb {

struct Contraction {
    int16_t p1, r1, p2, r2, p3, r3;  // signed planes, rings
    int32_t rp;                      // ring-triple index (calCR order)
    int32_t bin;                     // flat (K1_bin, T_bin, MU_bin)
};

struct PlanData {
    bb_geometry g{};
    int n_k1bin = 0, nbins = 0;
    std::vector<double> low, high, avg;      // ring edges
    std::vector<int32_t> nlo, nhi;           // ring r holds ia^2+ja^2 in [nlo, nhi)
    std::vector<int32_t> tri;                // rcount*3 (r1, r2, r3)
    std::vector<Contraction> contr;          // reference enumeration order
    std::vector<uint8_t> field_used;         // [Nmax*Nr]
    long long npairs = 0;
    int ngroups = 0;
    int nfields_used = 0;
};

// Sets the thread's error message and returns code.
int set_error(int code, const std::string &msg);

// Device-independent estimator layout (grid levels, field slots, work lists,
// binning records); built once per (plan, precision, options) and shared by
// every estimator of that key.
struct Layout;

}  // namespace bb

struct bb_plan {
    bb::PlanData d;
    std::mutex mu;   // guards the caches below
    std::map<std::string, std::shared_ptr<const bb::Layout>> layouts;
    std::vector<double> num;   // triangle counts (calCR), once computed
    bool have_num = false;
};
