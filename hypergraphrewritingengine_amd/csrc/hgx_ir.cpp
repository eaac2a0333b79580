// hgx_ir.cpp -- individualization-refinement canonical form on the host (see hgx_ir.hpp for
// the rules that fix the value and where the reference states them).
#include "hgx_ir.hpp"

#include <algorithm>
#include <cstring>
#include <numeric>

#include "hgx_device.hpp"

namespace hgx {
namespace ir {
namespace {

constexpr uint32_t kMaxDepth = 64;        // ir_core.hpp:50 with the host's escalation to it
constexpr uint32_t kMaxGenerators = 512;  // ir_core.hpp:66 (prunes only; never moves the value)
constexpr uint32_t kSposBits = 48;        // ir_core.hpp:28

// Lexicographic order of two ascending key runs, a proper prefix first (ir_core.hpp:212-219).
int cmp_runs(const uint64_t* a, size_t na, const uint64_t* b, size_t nb) {
    const size_t m = na < nb ? na : nb;
    for (size_t i = 0; i < m; ++i)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return na == nb ? 0 : (na < nb ? -1 : 1);
}

// An ordered partition: cells are named by ids 0..ncells-1; lab holds each cell's vertices
// contiguously from cstart.
struct Part {
    std::vector<uint32_t> lab, pos, cell, cstart, clen;
    uint32_t ncells = 0;
    void resize(uint32_t n) {
        lab.assign(n, 0); pos.assign(n, 0); cell.assign(n, 0);
        cstart.assign(n, 0); clen.assign(n, 0);
    }
};

class Canon {
public:
    Canon(uint32_t n, uint32_t ne, const uint8_t* ar, const uint32_t* eoff, const uint32_t* ev)
        : n_(n), ne_(ne), ar_(ar), eoff_(eoff), ev_(ev) {
        // occurrences of every vertex, in (edge, position) order
        occ_off_.assign(n + 1, 0);
        for (uint32_t e = 0; e < ne; ++e)
            for (uint32_t p = 0; p < ar[e]; ++p) ++occ_off_[ev[eoff[e] + p] + 1];
        for (uint32_t v = 0; v < n; ++v) occ_off_[v + 1] += occ_off_[v];
        occ_e_.resize(occ_off_[n]);
        occ_p_.resize(occ_off_[n]);
        std::vector<uint32_t> cur(occ_off_.begin(), occ_off_.end() - 1);
        for (uint32_t e = 0; e < ne; ++e)
            for (uint32_t p = 0; p < ar[e]; ++p) {
                const uint32_t w = cur[ev[eoff[e] + p]]++;
                occ_e_[w] = e;
                occ_p_[w] = p;
            }
        form_words_ = ne + occ_off_[n];
        stamp_.assign(ne, 0);
        mark_.assign(n, 0);
        sig_.resize(n);
    }

    // The hash; false when a path needs more individualizations than the reference allows.
    bool run(uint64_t& hash, std::vector<uint32_t>* form_out, uint32_t* rank_out) {
        Part root;
        initial_partition(root);
        refine(root);
        std::vector<uint32_t> path(kMaxDepth, 0);
        if (!search(root, 0, path)) return false;
        hash = fnv(FNV_OFFSET, n_);
        for (uint32_t w = 0, e = 0; e < ne_; ++e) {
            const uint32_t a = best_[w++];
            for (uint32_t k = 0; k < a; ++k) hash = fnv(hash, best_[w++]);
            hash = fnv(hash, 0xDEADBEEFull);
        }
        if (form_out) *form_out = best_;
        if (rank_out)
            for (uint32_t i = 0; i < ne_; ++i) rank_out[best_order_[i]] = i;
        return true;
    }

private:
    uint32_t n_, ne_;
    const uint8_t* ar_;
    const uint32_t *eoff_, *ev_;
    std::vector<uint32_t> occ_off_, occ_e_, occ_p_;
    uint32_t form_words_ = 0;
    std::vector<uint32_t> stamp_;
    uint32_t epoch_ = 0;
    std::vector<uint8_t> mark_;
    std::vector<std::vector<uint64_t>> sig_;
    std::vector<uint32_t> best_, best_order_, first_, first_lab_;
    bool have_best_ = false, have_first_ = false;
    std::vector<std::vector<uint32_t>> gens_;

    uint32_t vtx(uint32_t e, uint32_t p) const { return ev_[eoff_[e] + p]; }

    // Cells of equal degree signature, ascending by signature (ir_core.hpp:256-303).
    void initial_partition(Part& P) {
        P.resize(n_);
        std::vector<std::vector<uint64_t>> deg(n_);
        for (uint32_t v = 0; v < n_; ++v) {
            for (uint32_t o = occ_off_[v]; o < occ_off_[v + 1]; ++o)
                deg[v].push_back(((uint64_t)ar_[occ_e_[o]] << 32) | occ_p_[o]);
            std::sort(deg[v].begin(), deg[v].end());
        }
        std::vector<uint32_t> order(n_);
        std::iota(order.begin(), order.end(), 0u);
        std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
            const int c = cmp_runs(deg[a].data(), deg[a].size(), deg[b].data(), deg[b].size());
            return c != 0 ? c < 0 : a < b;
        });
        P.ncells = 0;
        for (uint32_t i = 0; i < n_;) {
            const uint32_t c = P.ncells++;
            uint32_t j = i;
            while (j < n_ && cmp_runs(deg[order[i]].data(), deg[order[i]].size(), deg[order[j]].data(),
                                      deg[order[j]].size()) == 0) {
                P.lab[j] = order[j]; P.pos[order[j]] = j; P.cell[order[j]] = c;
                ++j;
            }
            P.cstart[c] = i;
            P.clen[c] = j - i;
            i = j;
        }
    }

    // To the coarsest equitable partition finer than P (ir_core.hpp:311-511).
    void refine(Part& P) {
        const uint32_t words = n_ / 64 + 1;
        std::vector<uint64_t> queued(words, 0);
        for (uint32_t c = 0; c < P.ncells; ++c) queued[c >> 6] |= 1ull << (c & 63);
        std::vector<uint32_t> inc, touched, staging;
        while (P.ncells < n_) {
            uint32_t S = 0xFFFFFFFFu;
            for (uint32_t w = 0; w < words; ++w)
                if (queued[w]) {
                    const uint32_t b = (uint32_t)__builtin_ctzll(queued[w]);
                    queued[w] &= queued[w] - 1;
                    S = w * 64 + b;
                    break;
                }
            if (S == 0xFFFFFFFFu) break;
            if (S >= P.ncells) continue;
            // the edges on S's vertices, each once
            ++epoch_;
            inc.clear();
            for (uint32_t k = 0; k < P.clen[S]; ++k) {
                const uint32_t s = P.lab[P.cstart[S] + k];
                for (uint32_t o = occ_off_[s]; o < occ_off_[s + 1]; ++o)
                    if (stamp_[occ_e_[o]] != epoch_) { stamp_[occ_e_[o]] = epoch_; inc.push_back(occ_e_[o]); }
            }
            if (inc.empty()) continue;
            // every vertex on those edges: its multiset of (arity, position, splitter positions)
            touched.clear();
            for (uint32_t e : inc)
                for (uint32_t p = 0; p < ar_[e]; ++p) {
                    const uint32_t u = vtx(e, p);
                    if (!mark_[u]) { mark_[u] = 1; touched.push_back(u); sig_[u].clear(); }
                }
            for (uint32_t e : inc) {
                const uint32_t a = ar_[e];
                uint64_t spos = 0;
                for (uint32_t p = 0; p < a && p < kSposBits; ++p)
                    if (P.cell[vtx(e, p)] == S) spos |= 1ull << p;
                for (uint32_t p = 0; p < a; ++p)
                    sig_[vtx(e, p)].push_back(((uint64_t)(a & 0xFF) << 56) | ((uint64_t)(p & 0xFF) << 48) | spos);
            }
            for (uint32_t u : touched) std::sort(sig_[u].begin(), sig_[u].end());
            std::sort(touched.begin(), touched.end(), [&](uint32_t a, uint32_t b) {
                if (P.cell[a] != P.cell[b]) return P.cell[a] < P.cell[b];
                return cmp_runs(sig_[a].data(), sig_[a].size(), sig_[b].data(), sig_[b].size()) < 0;
            });
            // split each touched cell, in ascending cell id
            for (size_t i = 0; i < touched.size();) {
                const uint32_t C = P.cell[touched[i]];
                size_t j = i;
                while (j < touched.size() && P.cell[touched[j]] == C) ++j;
                std::vector<size_t> gb;                         // group starts in touched
                for (size_t k = i; k < j;) {
                    gb.push_back(k);
                    size_t m = k + 1;
                    while (m < j && cmp_runs(sig_[touched[m]].data(), sig_[touched[m]].size(),
                                             sig_[touched[k]].data(), sig_[touched[k]].size()) == 0) ++m;
                    k = m;
                }
                gb.push_back(j);
                const uint32_t groups = (uint32_t)gb.size() - 1;
                const uint32_t leftover = P.clen[C] - (uint32_t)(j - i);
                if (groups + (leftover ? 1u : 0u) > 1) {
                    const uint32_t cs = P.cstart[C];
                    staging.clear();
                    for (uint32_t k = 0; k < P.clen[C]; ++k)
                        if (!mark_[P.lab[cs + k]]) staging.push_back(P.lab[cs + k]);
                    for (size_t k = i; k < j; ++k) staging.push_back(touched[k]);
                    for (uint32_t k = 0; k < P.clen[C]; ++k) { P.lab[cs + k] = staging[k]; P.pos[staging[k]] = cs + k; }
                    // the largest piece keeps C (the untouched piece first, then groups, first wins)
                    uint32_t best_len = leftover, best_off = 0;
                    for (uint32_t g = 0, off = leftover; g < groups; ++g) {
                        const uint32_t len = (uint32_t)(gb[g + 1] - gb[g]);
                        if (len > best_len) { best_len = len; best_off = off; }
                        off += len;
                    }
                    auto assign = [&](uint32_t off, uint32_t len) {
                        if (!len) return;
                        uint32_t id = C;
                        if (!(off == best_off && len == best_len)) {
                            id = P.ncells++;
                            queued[id >> 6] |= 1ull << (id & 63);
                        }
                        P.cstart[id] = cs + off;
                        P.clen[id] = len;
                        for (uint32_t t = 0; t < len; ++t) P.cell[P.lab[cs + off + t]] = id;
                    };
                    assign(0, leftover);
                    for (uint32_t g = 0, off = leftover; g < groups; ++g) {
                        const uint32_t len = (uint32_t)(gb[g + 1] - gb[g]);
                        assign(off, len);
                        off += len;
                    }
                }
                for (size_t k = i; k < j; ++k) mark_[touched[k]] = 0;
                i = j;
            }
        }
    }

    // The target cell's vertex v singled out in front of the rest; every cell renumbered in
    // id order (ir_core.hpp:536-564).
    void individualize(const Part& src, Part& dst, uint32_t target, uint32_t v) const {
        dst.resize(n_);
        uint32_t w = 0, nc = 0;
        for (uint32_t c = 0; c < src.ncells; ++c) {
            const uint32_t cs = src.cstart[c], cl = src.clen[c];
            if (c == target) {
                const uint32_t id = nc++;
                dst.cstart[id] = w; dst.clen[id] = 1;
                dst.lab[w] = v; dst.pos[v] = w; dst.cell[v] = id; ++w;
                if (cl > 1) {
                    const uint32_t rid = nc++;
                    dst.cstart[rid] = w;
                    for (uint32_t k = 0; k < cl; ++k) {
                        const uint32_t u = src.lab[cs + k];
                        if (u == v) continue;
                        dst.lab[w] = u; dst.pos[u] = w; dst.cell[u] = rid; ++w;
                    }
                    dst.clen[rid] = w - dst.cstart[rid];
                }
            } else {
                const uint32_t id = nc++;
                dst.cstart[id] = w;
                for (uint32_t k = 0; k < cl; ++k) {
                    const uint32_t u = src.lab[cs + k];
                    dst.lab[w] = u; dst.pos[u] = w; dst.cell[u] = id; ++w;
                }
                dst.clen[id] = w - dst.cstart[id];
            }
        }
        dst.ncells = nc;
    }

    // A discrete partition: its form, the best so far, and an automorphism when it equals the
    // first leaf's form (ir_core.hpp:661-694).
    void leaf(const Part& P) {
        const std::vector<uint32_t>& lab = P.cell;            // vertex -> its singleton's id
        std::vector<uint32_t> order(ne_);
        std::iota(order.begin(), order.end(), 0u);
        std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
            const uint32_t la = ar_[a], lb = ar_[b], m = la < lb ? la : lb;
            for (uint32_t k = 0; k < m; ++k) {
                const uint32_t x = lab[vtx(a, k)], y = lab[vtx(b, k)];
                if (x != y) return x < y;
            }
            return la != lb ? la < lb : a < b;
        });
        std::vector<uint32_t> form;
        form.reserve(form_words_);
        for (uint32_t e : order) {
            form.push_back(ar_[e]);
            for (uint32_t k = 0; k < ar_[e]; ++k) form.push_back(lab[vtx(e, k)]);
        }
        if (!have_best_ || form < best_) { best_ = form; best_order_ = order; have_best_ = true; }
        if (!have_first_) {
            first_ = form;
            first_lab_ = lab;
            have_first_ = true;
        } else if (form == first_ && gens_.size() < kMaxGenerators) {
            std::vector<uint32_t> inv(n_), g(n_);
            for (uint32_t x = 0; x < n_; ++x) inv[lab[x]] = x;
            bool identity = true;
            for (uint32_t u = 0; u < n_; ++u) { g[u] = inv[first_lab_[u]]; identity &= g[u] == u; }
            if (!identity) gens_.push_back(std::move(g));
        }
    }

    // Depth-first over the individualizations of P (refined, at depth d); false when a path
    // would need more than kMaxDepth - 1 of them (ir_core.hpp:863-947).
    bool search(const Part& P, uint32_t d, std::vector<uint32_t>& path) {
        uint32_t t = 0;
        while (t < P.ncells && P.clen[t] < 2) ++t;
        if (t >= P.ncells) { leaf(P); return true; }
        std::vector<uint32_t> cell(P.lab.begin() + P.cstart[t], P.lab.begin() + P.cstart[t] + P.clen[t]);
        std::sort(cell.begin(), cell.end());
        std::vector<uint8_t> covered(cell.size(), 0);
        for (size_t ci = 0; ci < cell.size(); ++ci) {
            if (covered[ci]) continue;
            const uint32_t v = cell[ci];
            if (d + 1 >= kMaxDepth) return false;
            path[d] = v;
            Part child;
            individualize(P, child, t, v);
            refine(child);
            if (!search(child, d + 1, path)) return false;
            // every cell vertex automorphic to v under the automorphisms fixing the path above
            std::vector<uint32_t> uf(n_);
            std::iota(uf.begin(), uf.end(), 0u);
            auto find = [&](uint32_t x) {
                while (uf[x] != x) { uf[x] = uf[uf[x]]; x = uf[x]; }
                return x;
            };
            for (const auto& g : gens_) {
                bool fixes = true;
                for (uint32_t k = 0; k < d && fixes; ++k) fixes = g[path[k]] == path[k];
                if (!fixes) continue;
                for (uint32_t i = 0; i < n_; ++i) {
                    const uint32_t a = find(i), b = find(g[i]);
                    if (a != b) uf[a] = b;
                }
            }
            const uint32_t rv = find(v);
            for (size_t k = 0; k < cell.size(); ++k)
                if (find(cell[k]) == rv) covered[k] = 1;
        }
        return true;
    }
};

}  // namespace

uint64_t state_hash(const uint32_t* edge_off, const uint32_t* verts, uint32_t n_edges,
                    std::vector<uint32_t>* form, uint32_t* rank) {
    if (form) form->clear();
    if (rank) std::fill(rank, rank + n_edges, 0u);
    if (n_edges == 0) return 0;
    // local numbering: distinct labels ascending (ir_canonicalization.cpp:42-58)
    const uint32_t base = edge_off[0], occ = edge_off[n_edges] - base;
    std::vector<uint32_t> sorted(verts + base, verts + base + occ);
    std::sort(sorted.begin(), sorted.end());
    sorted.erase(std::unique(sorted.begin(), sorted.end()), sorted.end());
    std::vector<uint8_t> ar(n_edges);
    std::vector<uint32_t> eoff(n_edges), ev(occ);
    for (uint32_t e = 0; e < n_edges; ++e) {
        ar[e] = (uint8_t)(edge_off[e + 1] - edge_off[e]);
        eoff[e] = edge_off[e] - base;
    }
    for (uint32_t i = 0; i < occ; ++i)
        ev[i] = (uint32_t)(std::lower_bound(sorted.begin(), sorted.end(), verts[base + i]) - sorted.begin());
    Canon c((uint32_t)sorted.size(), n_edges, ar.data(), eoff.data(), ev.data());
    uint64_t h = 0;
    if (!c.run(h, form, rank)) {
        if (form) form->clear();
        if (rank) std::fill(rank, rank + n_edges, 0u);
        return 0;
    }
    return h;
}

}  // namespace ir
}  // namespace hgx

// ---- C ABI ---------------------------------------------------------------------------------

#include <atomic>
#include <thread>

#include "../../include/hgx.h"

namespace {

bool state_ok(const uint32_t* edge_off, uint32_t e0, uint32_t e1) {
    for (uint32_t e = e0; e < e1; ++e)
        if (edge_off[e + 1] < edge_off[e] || edge_off[e + 1] - edge_off[e] < 1 || edge_off[e + 1] - edge_off[e] > 16)
            return false;
    return true;
}

}  // namespace

extern "C" int hgx_ir_hash_batch(hgx_ctx* ctx, const uint64_t* state_offsets, uint32_t n_states,
                                 const uint32_t* edge_offsets, const uint32_t* vertices, uint64_t* out_hash) {
    return hgx_ir_rank_batch(ctx, state_offsets, n_states, edge_offsets, vertices, out_hash, nullptr);
}

extern "C" int hgx_ir_rank_batch(hgx_ctx* ctx, const uint64_t* state_offsets, uint32_t n_states,
                                 const uint32_t* edge_offsets, const uint32_t* vertices, uint64_t* out_hash,
                                 uint32_t* out_rank) {
    (void)ctx;
    if (n_states == 0) return HGX_OK;
    if (!state_offsets || !out_hash) return HGX_ERR_INVALID;
    for (uint32_t s = 0; s < n_states; ++s) {
        if (state_offsets[s + 1] < state_offsets[s] || state_offsets[s + 1] > 0xFFFFFFFFull) return HGX_ERR_INVALID;
        if (state_offsets[s + 1] > state_offsets[s] && (!edge_offsets || !vertices)) return HGX_ERR_INVALID;
        if (!state_ok(edge_offsets, (uint32_t)state_offsets[s], (uint32_t)state_offsets[s + 1])) return HGX_ERR_INVALID;
    }
    std::atomic<uint32_t> next{0};
    auto work = [&]() {
        for (uint32_t s; (s = next.fetch_add(1)) < n_states;) {
            const uint32_t e0 = (uint32_t)state_offsets[s], ne = (uint32_t)(state_offsets[s + 1] - e0);
            out_hash[s] = ne ? hgx::ir::state_hash(edge_offsets + e0, vertices, ne, nullptr,
                                                   out_rank ? out_rank + (e0 - (uint32_t)state_offsets[0]) : nullptr)
                             : 0;
        }
    };
    const uint32_t hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const uint32_t nt = std::min(hw, n_states);
    std::vector<std::thread> pool;
    for (uint32_t t = 1; t < nt; ++t) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    return HGX_OK;
}

extern "C" int hgx_ir_canonical_form(const uint32_t* edge_offsets, const uint32_t* vertices, uint32_t n_edges,
                                     uint32_t* out_edge_offsets, uint32_t* out_vertices, uint64_t* out_hash) {
    if (!out_hash || (n_edges && (!edge_offsets || !vertices || !out_edge_offsets || !out_vertices)))
        return HGX_ERR_INVALID;
    if (n_edges && !state_ok(edge_offsets, 0, n_edges)) return HGX_ERR_INVALID;
    std::vector<uint32_t> form;
    *out_hash = hgx::ir::state_hash(edge_offsets, vertices, n_edges, &form);
    if (!n_edges) return HGX_OK;
    if (form.empty()) return HGX_ERR_UNSUPPORTED;          // the search gave up (hash 0)
    out_edge_offsets[0] = 0;
    uint32_t w = 0, o = 0;
    for (uint32_t e = 0; e < n_edges; ++e) {
        const uint32_t a = form[w++];
        for (uint32_t k = 0; k < a; ++k) out_vertices[o++] = form[w++];
        out_edge_offsets[e + 1] = o;
    }
    return HGX_OK;
}
