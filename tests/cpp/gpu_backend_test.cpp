// gpu_backend_test.cpp -- a caller of run_gpu_evolution (include/hgx_gpu_backend.hpp) shaped like
// the paclet's run_gpu_job (hypergraph_ffi.cpp:739-742): it fills a GpuJob and writes the WXF
// result of each case to OUTDIR/case<i>.wxf; a refused job prints "refused <i>: <reason>".
//   gpu_backend_test OUTDIR
#include <cstdio>
#include <string>
#include <vector>

#include "hgx_gpu_backend.hpp"

using Edges = std::vector<std::vector<int64_t>>;
using Rules = std::vector<std::pair<std::string, std::vector<Edges>>>;

static bool write(const std::string& path, const std::vector<uint8_t>& b) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(b.data(), 1, b.size(), f) == b.size();
    std::fclose(f);
    return ok;
}

int main(int argc, char** argv) {
    if (argc < 2) { std::fprintf(stderr, "usage: gpu_backend_test OUTDIR\n"); return 2; }
    const std::string out = argv[1];
    const Rules rules = {{"Rule1", {{{0, 1}, {1, 2}}, {{0, 1}, {1, 2}, {2, 0}}}}};
    const std::vector<Edges> init = {{{1, 2}, {2, 3}, {3, 1}}};
    const Rules rules2 = {{"a", {{{0, 1}, {0, 2}}, {{0, 2}, {0, 3}, {1, 3}, {2, 3}}}},
                          {"b", {{{0, 0}}, {{0, 1}, {1, 0}}}}};
    const std::vector<Edges> init2 = {{{1, 1}, {1, 1}}, {{5, 6}, {6, 5}}};
    hgx::HostBridge host;
    int progress = 0;
    host.progress = [&](const std::string&) { ++progress; };
    int rc = 0;
    for (int i = 0; i < 7; ++i) {
        hgx::GpuJob job{i < 2 ? rules : rules2, i < 2 ? init : init2};
        job.steps = 3;
        if (i == 1) {
            job.state_canon_mode = hgx::GpuJob::StateCanonCode::kNone;
            job.event_canon_mode = hgx::GpuJob::EventCanonCode::kAutomatic;
            job.include_canonical_hashes = true;
            job.transitive_reduction = false;
        }
        if (i == 3) {
            job.exploration_probability = 0.5;
            job.transition_rate = 0.75;
            job.random_seed = 12345;
        }
        if (i == 5) {
            job.explore_from_canonical_states_only = true;
            job.quotient_initial_states = true;
            job.event_canon_mode = hgx::GpuJob::EventCanonCode::kAutomatic;
        }
        if (i == 6) job.show_genesis_events = true;
        if (i == 4) {
            job.state_canon_mode = hgx::GpuJob::StateCanonCode::kAutomatic;
            job.event_canon_mode = hgx::GpuJob::EventCanonCode::kFull;
        }
        try {
            const std::vector<uint8_t> b = hgx::run_gpu_evolution(job, host);
            if (!write(out + "/case" + std::to_string(i) + ".wxf", b)) rc = 3;
        } catch (const std::exception& e) {
            std::printf("refused %d: %s\n", i, e.what());
        }
    }
    std::printf("progress %d\n", progress);
    return rc;
}
