// tune.hip -- autotuner of one model's launch shapes: levels per work-group
// of the two k-marching kernels (tendency; friction + Adams-Bashforth), the
// form of the Thomas solve (whole columns in registers, or columns per
// wavefront of the LDS form) and, in the mixed form, the radix plan of each
// FP32 transform axis.  Each candidate is timed with
// HIP events on the live model; the launches timed do not change the model's
// state (the tendency writes the slot the next step overwrites anyway, the
// friction kernel writes the second buffers, Thomas runs on a scratch copy of
// the spectral array).  Results do not depend on the shapes.  The choice is
// cached per (stage, nx x ny x nz, boundary, precision) in a text file of
// "stage shape value" lines, the same format as the ensemble's (whose
// stages carry an "ens_" prefix: a batched launch is timed differently).
#include "ps3d_internal.h"

#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>
#include <vector>

namespace ps3d {

namespace {

constexpr size_t kFlushBytes = size_t(512) << 20;   // > 8 x 4 MiB L2 + 256 MiB Infinity Cache

std::string shape_of(const Model &m)
{
    std::ostringstream o;
    o << m.g.nx << "x" << m.g.ny << "x" << m.g.nz << (m.prm.enable_vertical_boundaries ? ":lid" : ":periodic")
      << (m.mixed ? ":mixed" : ":fp64");
    return o.str();
}

template <class F>
float time_it(Model &m, int reps, F &&launch)
{
    hipEvent_t a, b;
    PS3D_HIP_CHECK(hipEventCreate(&a));
    PS3D_HIP_CHECK(hipEventCreate(&b));
    float ms = 0.f;
    try {
        launch();   // warm-up
        PS3D_HIP_CHECK(hipEventRecord(a, m.stream));
        for (int r = 0; r < reps; ++r) launch();
        PS3D_HIP_CHECK(hipEventRecord(b, m.stream));
        PS3D_HIP_CHECK(hipEventSynchronize(b));
        PS3D_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
    } catch (...) {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        throw;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms;
}

// Time of reps launches, each after the caches were flushed by writing a
// buffer larger than the L2s and the Infinity Cache together: for kernels
// that read static tables, which back-to-back launches would find cached
// while the step's other kernels evict them.
template <class F>
float time_it_cold(Model &m, int reps, void *flush, size_t flush_bytes, F &&launch)
{
    hipEvent_t a, b;
    PS3D_HIP_CHECK(hipEventCreate(&a));
    PS3D_HIP_CHECK(hipEventCreate(&b));
    float total = 0.f;
    try {
        launch();   // warm-up
        for (int r = 0; r < reps; ++r) {
            PS3D_HIP_CHECK(hipMemsetAsync(flush, r & 0xff, flush_bytes, m.stream));
            PS3D_HIP_CHECK(hipEventRecord(a, m.stream));
            launch();
            PS3D_HIP_CHECK(hipEventRecord(b, m.stream));
            PS3D_HIP_CHECK(hipEventSynchronize(b));
            float ms = 0.f;
            PS3D_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
            total += ms;
        }
    } catch (...) {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        throw;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return total;
}

}  // namespace

// "stage shape value" lines of a tuning cache; a missing file is an empty
// cache, a malformed line an error (so no entry is silently dropped)
std::map<std::string, int> tune_cache_read(const char *path)
{
    std::map<std::string, int> cache;
    if (!path || !*path) return cache;
    std::ifstream in(path);
    if (!in.is_open()) return cache;
    std::string line;
    int lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
        std::istringstream ls(line);
        std::string st, sh, extra;
        int v = 0;
        PS3D_REQUIRE(bool(ls >> st >> sh >> v) && !(ls >> extra),
                     std::string("tune cache ") + path + ": malformed line " + std::to_string(lineno));
        cache[st + " " + sh] = v;
    }
    PS3D_REQUIRE(!in.bad(), std::string("tune cache ") + path + ": read failed");
    return cache;
}

// written to a temporary file beside it and renamed over it, so a reader
// never sees a partly written cache
void tune_cache_write(const char *path, const std::map<std::string, int> &cache)
{
    if (!path || !*path) return;
    const std::string tmp = std::string(path) + ".tmp" + std::to_string((long long)getpid());
    {
        std::ofstream out(tmp, std::ios::trunc);
        PS3D_REQUIRE(out.is_open(), "tune cache: cannot write " + tmp);
        for (auto &kv : cache) out << kv.first << " " << kv.second << "\n";
        out.flush();
        PS3D_REQUIRE(bool(out), "tune cache: write failed: " + tmp);
    }
    if (std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        throw Error(PS3D_ERR_ARG, std::string("tune cache: cannot rename ") + tmp + " to " + path);
    }
}

void tune_model(Model &m, int next_tau, const char *cache_path, int reps)
{
    PS3D_REQUIRE(next_tau >= 1 && next_tau <= 3, "tune: next_tau must be 1..3");
    for (int f = 0; f < 4; ++f) PS3D_REQUIRE(m.cur[f] == 0, "tune: friction results not yet folded in by ab_update");
    reps = reps < 1 ? 1 : reps;
    std::map<std::string, int> cache = tune_cache_read(cache_path);
    const std::string sh = shape_of(m);
    const int nz = m.g.nz;
    std::vector<int> levels;
    for (int kc = 2; kc < nz; kc *= 2) levels.push_back(kc);
    levels.push_back(nz);
    const int t1 = next_tau % 3 + 1, t2 = t1 % 3 + 1;   // any two other slots

    // flush: scratch for time_it_cold (null: back-to-back timing)
    auto pick = [&](const std::string &stage, const std::vector<int> &cand, int &field, int nrep, void *flush,
                    auto &&launch) {
        const std::string key = stage + " " + sh;
        auto hit = cache.find(key);
        if (hit != cache.end()) {
            field = hit->second;
            return;
        }
        // three rounds over the candidates, interleaved, and the median of
        // each candidate's rounds: the device's clocks drift over a tuning
        // run by more than the few per cent that separate close candidates
        constexpr int kRounds = 3;
        std::vector<std::vector<float>> t(cand.size());
        for (int round = 0; round < kRounds; ++round)
            for (size_t q = 0; q < cand.size(); ++q) {
                field = cand[q];
                t[q].push_back(flush ? time_it_cold(m, nrep, flush, kFlushBytes, launch) : time_it(m, nrep, launch));
            }
        int best = cand[0];
        float best_ms = 1e30f;
        for (size_t q = 0; q < cand.size(); ++q) {
            std::sort(t[q].begin(), t[q].end());
            if (t[q][kRounds / 2] < best_ms) {
                best_ms = t[q][kRounds / 2];
                best = cand[q];
            }
        }
        field = best;
        cache[key] = best;
    };

    if (fused_tendency_active(m)) {
        // the fused kernel writes slot next_tau (overwritten by the next
        // step) and the second buffers
        pick("tend_friction_ab", levels, m.tfric_kc, reps, nullptr,
             [&] { launch_friction_ab_raw(m, 3, next_tau, t1, t2, true); });
    } else {
        pick("tendency", levels, m.tend_kc, reps, nullptr, [&] {
            if (m.mixed) launch_tendency_mixed(m, next_tau, m.pint);
            else launch_tendency(m, next_tau);
        });
        if (m.mixed || (fused_friction_ok(m) && m.fuse_friction))
            pick("friction_ab", levels, m.fric_kc, reps, nullptr,
                 [&] { launch_friction_ab_raw(m, 3, next_tau, t1, t2); });
    }
    if (m.mixed) {
        // radix plans of the FP32 transforms, on scratch arrays: the forward
        // x pass reads the state and writes the scratch spectrum, the other
        // passes work on that
        float2 *fs = nullptr;
        float *pd = nullptr;
        try {
            const int lens[3] = {m.g.nx / 2, m.g.ny, m.g.nz};
            const char *names[3] = {"fft_plan_x", "fft_plan_y", "fft_plan_z"};
            for (int ax = 0; ax < 3; ++ax) {
                if (ax == 2 && m.prm.enable_vertical_boundaries) continue;
                if (!fft_plan_choice(lens[ax])) continue;
                if (!fs) {
                    PS3D_HIP_CHECK(hipMalloc(&fs, sizeof(float2) * (size_t)m.g.S));
                    PS3D_HIP_CHECK(hipMalloc(&pd, sizeof(float) * (size_t)m.g.H));
                }
                pick(names[ax], {0, 1}, m.fft_plan[ax], reps, nullptr, [&] {
                    fft_r2c_x_f(m, m.fh[0], m.fh[1], m.fh[2], fs);
                    if (ax == 0) fft_c2r_x_f(m, fs, pd, 1.0);
                    if (ax >= 1) fft_c2c_y_f(m, fs, -1, 0.0, 0.0);
                    if (ax == 2) fft_z_periodic_solve_f(m, fs);
                    if (ax >= 1) fft_c2c_y_f(m, fs, +1, 0.0, 0.0);
                });
            }
        } catch (...) {
            if (fs) (void)hipFree(fs);
            if (pd) (void)hipFree(pd);
            throw;
        }
        if (fs) {
            PS3D_HIP_CHECK(hipStreamSynchronize(m.stream));
            PS3D_HIP_CHECK(hipFree(fs));
            PS3D_HIP_CHECK(hipFree(pd));
        }
    }
    if (m.prm.enable_vertical_boundaries && !cache.count("thomas " + sh)) {
        // forms (launch_thomas, launch_thomas_f): 1 whole columns in
        // registers (k_thomas_reg) and 2 the real and imaginary parts on two
        // lanes (k_thomas_split), where nz is one of their compiled-in
        // lengths; FP64: 16, 32, 64 columns per wave on the fly, whose LDS
        // keeps several work-groups on a CU (64 KiB; wider groups, which a
        // back-to-back timing can prefer, measured slower in the step), and
        // -16, -32, -64 the factor tables; mixed: 32 columns on the fly, and
        // -64 the factor tables with FP64 d' in the scratch array where the
        // model has one
        const bool mx = m.mixed;
        const size_t esz = mx ? sizeof(float2) : sizeof(double2);
        const uint64_t fsb = esz * (uint64_t)m.g.S;
        std::vector<int> forms;
        if (thomas_reg_ok(m.g.nz, fsb)) forms.push_back(1);
        if (thomas_split_ok(m.g.nz, fsb)) forms.push_back(2);
        if (mx) {
            if ((sizeof(double2) + sizeof(double)) * 32 * (size_t)m.g.nz <= 160 * 1024) forms.push_back(32);
            if (m.dp64) forms.push_back(-64);
        } else {
            for (int c : {32, 16, 64})
                if ((sizeof(double2) + sizeof(double)) * (size_t)c * m.g.nz <= 64 * 1024) forms.push_back(c);
            for (int c : {-64, -32, -16}) forms.push_back(c);
        }
        void *scratch = nullptr, *flush = nullptr;
        try {
            PS3D_HIP_CHECK(hipMalloc(&scratch, esz * (size_t)m.g.S));
            PS3D_HIP_CHECK(hipMalloc(&flush, kFlushBytes));
            PS3D_HIP_CHECK(hipMemcpyAsync(scratch, mx ? (const void *)m.FSf : (const void *)m.FS,
                                          esz * (size_t)m.g.S, hipMemcpyDeviceToDevice, m.stream));
            // the table form reads per-column factors that stay cached when
            // launched back to back: timed with cold caches, as in the step
            int form = 0;
            pick("thomas", forms, form, 2 * reps, flush, [&] {
                m.thomas_cpw = form;
                if (mx) launch_thomas_f(m, static_cast<float2 *>(scratch));
                else launch_thomas(m, static_cast<double2 *>(scratch));
            });
            m.thomas_cpw = form;
        } catch (...) {
            if (scratch) (void)hipFree(scratch);
            if (flush) (void)hipFree(flush);
            throw;
        }
        PS3D_HIP_CHECK(hipStreamSynchronize(m.stream));
        PS3D_HIP_CHECK(hipFree(scratch));
        PS3D_HIP_CHECK(hipFree(flush));
    } else if (m.prm.enable_vertical_boundaries) {
        m.thomas_cpw = cache["thomas " + sh];
    }
    tune_cache_write(cache_path, cache);
}

std::string describe_tuning(const Model &m)
{
    std::ostringstream o;
    const std::string sh = shape_of(m);
    if (fused_tendency_active(m)) {
        o << "tend_friction_ab " << sh << " " << m.tfric_kc << "\n";
    } else {
        o << "tendency " << sh << " " << m.tend_kc << "\n";
        o << "friction_ab " << sh << " " << m.fric_kc << "\n";
    }
    o << "thomas " << sh << " " << m.thomas_cpw << "\n";
    if (m.mixed) {
        const char *names[3] = {"fft_plan_x", "fft_plan_y", "fft_plan_z"};
        for (int ax = 0; ax < 3; ++ax) o << names[ax] << " " << sh << " " << m.fft_plan[ax] << "\n";
    }
    return o.str();
}

}  // namespace ps3d
