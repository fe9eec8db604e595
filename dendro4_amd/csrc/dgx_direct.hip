// dgx_direct.hip -- sparse FP64 direct solver of the coarsest multigrid level
// (nested dissection + multifrontal Gauss-Jordan sweeps), see dgx_direct.h.
//
// Reference counterpart: the coarse-grid PCLU of the DAMG V-cycle
// (bin/optionsFBM:57, src/omg/blockDiag.cpp:780-1047 and 1172-1308), a sparse
// LU of the coarsest level's matrix through PETSc.
#include "dgx_direct.h"
#include <cstdio>
#include <string>
#include <chrono>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <limits>
#include <numeric>
#include <thread>

#include "../../include/dendro_mi355x.h"

// ------------------------------------------------------------ host plan ----
namespace {

// Runs body(lo, hi, t) over [0, n) in up to 8 contiguous ranges on host
// threads; each range writes only its own outputs, so the result does not
// depend on the split.
template <typename F>
void parallel_ranges(size_t n, size_t grain, F body) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned t = (unsigned)std::min<size_t>(std::min(8u, hw), std::max<size_t>(1, n / grain));
    if (t <= 1) {
        body(size_t(0), n, 0u);
        return;
    }
    std::vector<std::thread> th;
    for (unsigned k = 1; k < t; ++k) th.emplace_back(body, n * k / t, n * (k + 1) / t, k);
    body(size_t(0), n / t, 0u);
    for (auto &x : th) x.join();
}

// Geometric nested dissection of the interior-node graph.  A part is split by
// a plane c_a = t: L = nodes with c_a < t, the separator S = nodes with
// c_a >= t that have a neighbour in L, R = the rest; no edge joins L and R.
// Candidate planes lie on the grid lines of the largest power-of-two spacing
// that cuts the part's extent into at least kBins pieces, so they follow the
// octree's own cell faces.  The plane minimises |S| / min(|L|, |R|) over the
// three axes with both sides holding at least 15% of the part, so cuts go
// through coarse regions of an adaptive mesh where they are cheap.  The counts
// come from histograms: #L(t) = #{c < t}, #S(t) = #{m < t} - #{c < t} where m is
// the smallest coordinate over a node and its neighbours in the part.
struct Dissection {
    static constexpr uint32_t kBins = 256;
    // parts at least this large split their two halves over two host threads,
    // down to kParDepth levels of the tree (at most 2^kParDepth threads)
    static constexpr size_t kParMin = 8192;
    static constexpr int kParDepth = 3;
    uint32_t ni = 0, leaf = 128;
    const std::vector<uint32_t> &roff, &rcol;  // interior graph (diagonal included)
    const std::vector<uint32_t> &xyz;          // 3 per interior node
    // stamp[v] = id of the part v was last placed in; parts split concurrently
    // are disjoint, and a neighbour in another part never carries this part's
    // id, so relaxed accesses suffice
    std::vector<std::atomic<uint32_t>> stamp;
    std::atomic<uint32_t> next_id{0};

    // A subtree: its fronts in creation order (children before parents), each
    // front's nodes and children (indices into this list), and its roots.
    struct Forest {
        std::vector<std::vector<uint32_t>> S, kids;
        std::vector<uint32_t> roots;
    };

    Dissection(const std::vector<uint32_t> &ro, const std::vector<uint32_t> &rc, const std::vector<uint32_t> &c,
               uint32_t n_int, uint32_t lf)
        : ni(n_int), leaf(lf), roff(ro), rcol(rc), xyz(c), stamp(n_int) {
        for (auto &a : stamp) a.store(0, std::memory_order_relaxed);
    }

    static Forest single(std::vector<uint32_t> &&P) {
        Forest f;
        f.S.push_back(std::move(P));
        f.kids.emplace_back();
        f.roots.push_back(0);
        return f;
    }

    // smallest coordinate along each axis over v and its neighbours in the part
    void low3(uint32_t v, uint32_t cur, uint32_t *m) const {
        m[0] = xyz[3 * (size_t)v];
        m[1] = xyz[3 * (size_t)v + 1];
        m[2] = xyz[3 * (size_t)v + 2];
        for (uint32_t k = roff[v]; k < roff[v + 1]; ++k) {
            const uint32_t w = rcol[k];
            if (stamp[w].load(std::memory_order_relaxed) != cur) continue;
            const uint32_t *c = &xyz[3 * (size_t)w];
            m[0] = std::min(m[0], c[0]);
            m[1] = std::min(m[1], c[1]);
            m[2] = std::min(m[2], c[2]);
        }
    }

    Forest run(std::vector<uint32_t> P, int depth = 0) {
        const size_t np = P.size();
        if (np == 0) return {};
        if (np <= leaf) return single(std::move(P));
        const uint32_t cur = next_id.fetch_add(1, std::memory_order_relaxed) + 1;
        for (uint32_t v : P) stamp[v].store(cur, std::memory_order_relaxed);
        std::vector<uint32_t> mlow(3 * np);
        uint32_t bmin[3] = {~0u, ~0u, ~0u}, bmax[3] = {0, 0, 0};
        for (size_t i = 0; i < np; ++i) {
            low3(P[i], cur, &mlow[3 * i]);
            for (int a = 0; a < 3; ++a) {
                const uint32_t c = xyz[3 * (size_t)P[i] + a];
                bmin[a] = std::min(bmin[a], c);
                bmax[a] = std::max(bmax[a], c);
            }
        }
        int best_a = -1;
        uint32_t best_t = 0;
        double best_cost = std::numeric_limits<double>::infinity();
        const size_t lo = std::max<size_t>(1, (size_t)(0.15 * (double)np));
        uint32_t hc[kBins + 1], hm[kBins + 1];
        for (int a = 0; a < 3; ++a) {
            if (bmax[a] == bmin[a]) continue;
            int sh = 0;  // grid spacing 2^sh: the extent spans at most kBins cells
            while (((bmax[a] >> sh) - (bmin[a] >> sh)) >= kBins) ++sh;
            const uint32_t base = bmin[a] >> sh;
            std::fill(hc, hc + kBins + 1, 0u);
            std::fill(hm, hm + kBins + 1, 0u);
            for (size_t i = 0; i < np; ++i) {
                ++hc[(xyz[3 * (size_t)P[i] + a] >> sh) - base];
                ++hm[(mlow[3 * i + a] >> sh) - base];
            }
            // plane t_b = (base + b) << sh: #{c < t_b} = sum of hc below b
            size_t nL = 0, nM = 0;
            for (uint32_t b = 1; b <= kBins; ++b) {
                nL += hc[b - 1];
                nM += hm[b - 1];
                if (nL < lo) continue;
                if (np - nL < lo) break;
                const size_t nS = nM - nL, nR = np - nM;
                if (nR < lo) continue;
                const double cost = (double)nS / (double)std::min(nL, nR);
                if (cost < best_cost) {
                    best_cost = cost;
                    best_a = a;
                    best_t = (base + b) << sh;
                }
            }
        }
        if (best_a < 0) return single(std::move(P));  // no admissible plane: one dense front
        std::vector<uint32_t> L, R, Sv;
        for (size_t i = 0; i < np; ++i) {
            const uint32_t v = P[i];
            if (xyz[3 * (size_t)v + best_a] < best_t) L.push_back(v);
            else if (mlow[3 * i + best_a] < best_t) Sv.push_back(v);
            else R.push_back(v);
        }
        std::vector<uint32_t>().swap(P);
        std::vector<uint32_t>().swap(mlow);
        Forest fl, fr;
        if (depth < kParDepth && L.size() >= kParMin && R.size() >= kParMin) {
            std::thread th([&] { fl = run(std::move(L), depth + 1); });
            fr = run(std::move(R), depth + 1);
            th.join();
        } else {
            fl = run(std::move(L), depth + 1);
            fr = run(std::move(R), depth + 1);
        }
        // the fronts of L, then those of R (renumbered), then the separator
        Forest out;
        const uint32_t shift = (uint32_t)fl.S.size();
        out.S = std::move(fl.S);
        out.kids = std::move(fl.kids);
        out.roots = std::move(fl.roots);
        for (size_t f = 0; f < fr.S.size(); ++f) {
            for (uint32_t &c : fr.kids[f]) c += shift;
            out.S.push_back(std::move(fr.S[f]));
            out.kids.push_back(std::move(fr.kids[f]));
        }
        for (uint32_t r : fr.roots) out.roots.push_back(r + shift);
        if (Sv.empty()) return out;  // the part was disconnected along the plane
        out.S.push_back(std::move(Sv));
        out.kids.push_back(std::move(out.roots));
        out.roots.assign(1, (uint32_t)(out.S.size() - 1));
        return out;
    }
};

}  // namespace

MfPlan mf_plan(uint32_t n, size_t E, const uint32_t *conn, const uint16_t *meta, const uint8_t *bdy,
               const uint32_t *xyz, uint64_t max_entries, uint32_t leaf, uint32_t dense) {
    MfPlan P;
    P.n = n;
    static const bool timing = [] {
        const char *e = getenv("DGX_SETUP_TIMING");
        return e && e[0] == '1';
    }();
    auto tp = std::chrono::steady_clock::now();
    std::string tlog;
    auto lap = [&](const char *what) {
        if (!timing) return;
        const auto now = std::chrono::steady_clock::now();
        char b[64];
        snprintf(b, sizeof b, " %s %.2f", what, std::chrono::duration<double, std::milli>(now - tp).count());
        tlog += b;
        tp = now;
    };
    if (E * 64 > 0xffffffffull) {
        dgx_set_error("coarse solver: %zu elements exceed the 32-bit contribution index", E);
        throw DgxError();
    }
    // interior numbering
    std::vector<int32_t> iid(n, -1);
    for (uint32_t i = 0; i < n; ++i)
        if (!bdy[i]) {
            iid[i] = (int32_t)P.inode.size();
            P.inode.push_back(i);
        }
    const uint32_t ni = (uint32_t)P.inode.size();
    P.ni = ni;
    // interior ids of each element's vertices (original vertex order), and the
    // incidences (interior node, element, vertex) in element order
    std::vector<int32_t> eid(8 * E);
    std::vector<uint32_t> ioff(ni + 1, 0);
    for (size_t e = 0; e < E; ++e)
        for (int j = 0; j < 8; ++j) {
            const int32_t v = iid[conn[(size_t)j * E + e]];
            eid[8 * e + j] = v;
            if (v >= 0) ++ioff[v + 1];
        }
    for (uint32_t i = 0; i < ni; ++i) ioff[i + 1] += ioff[i];
    std::vector<uint32_t> inc(ioff[ni]);  // e * 8 + vertex
    {
        std::vector<uint32_t> fill(ioff.begin(), ioff.end() - 1);
        for (size_t e = 0; e < E; ++e)
            for (uint32_t j = 0; j < 8; ++j) {
                const int32_t v = eid[8 * e + j];
                if (v >= 0) inc[fill[v]++] = (uint32_t)(e * 8 + j);
            }
    }
    // matrix pattern over the interior rows (diagonal included), columns in
    // the order the incident elements first reach them
    std::vector<uint32_t> roff(ni + 1, 0), rcol;
    {
        std::vector<std::vector<uint32_t>> part(8);
        parallel_ranges(ni, 4096, [&](size_t lo, size_t hi, unsigned t) {
            std::vector<uint32_t> mark(ni, 0xffffffffu), &out = part[t];
            out.reserve((hi - lo) * 27);
            for (uint32_t i = (uint32_t)lo; i < hi; ++i) {
                const size_t before = out.size();
                for (uint32_t q = ioff[i]; q < ioff[i + 1]; ++q) {
                    const int32_t *ids = &eid[(inc[q] >> 3) * 8];
                    for (int j = 0; j < 8; ++j) {
                        const int32_t w = ids[j];
                        if (w >= 0 && mark[w] != i) {
                            mark[w] = i;
                            out.push_back((uint32_t)w);
                        }
                    }
                }
                if (out.size() == before) out.push_back(i);  // an interior node of no element keeps its diagonal
                roff[i + 1] = (uint32_t)(out.size() - before);
            }
        });
        for (uint32_t i = 0; i < ni; ++i) roff[i + 1] += roff[i];
        rcol.reserve(roff[ni]);
        for (auto &pv : part) rcol.insert(rcol.end(), pv.begin(), pv.end());
    }
    const uint32_t nnz = roff[ni];
    lap("pattern");
    std::vector<uint32_t> cxyz(3 * (size_t)ni);
    for (uint32_t i = 0; i < ni; ++i)
        for (int a = 0; a < 3; ++a) cxyz[3 * (size_t)i + a] = xyz[3 * (size_t)P.inode[i] + a];

    // nested dissection; positions in front order (postorder), nodes in list order
    struct {
        std::vector<std::vector<uint32_t>> S, kids;
        std::vector<uint32_t> pos;
    } D;
    {
        Dissection nd(roff, rcol, cxyz, ni, ni <= dense ? std::max(ni, 1u) : leaf);
        std::vector<uint32_t> all(ni);
        std::iota(all.begin(), all.end(), 0u);
        Dissection::Forest fo = nd.run(std::move(all));
        D.S = std::move(fo.S);
        D.kids = std::move(fo.kids);
        D.pos.assign(ni, 0);
        uint32_t next_pos = 0;
        for (const auto &sep : D.S)
            for (uint32_t v : sep) D.pos[v] = next_pos++;
    }
    const uint32_t nf = (uint32_t)D.S.size();
    lap("dissection");
    std::vector<uint32_t> front_of(ni), sstart(nf), send(nf), at_pos(ni);
    for (uint32_t f = 0; f < nf; ++f) {
        sstart[f] = D.pos[D.S[f][0]];
        send[f] = sstart[f] + (uint32_t)D.S[f].size();
        for (uint32_t v : D.S[f]) {
            front_of[v] = f;
            at_pos[D.pos[v]] = v;
        }
    }
    // update sets, as elimination positions (ascending)
    std::vector<std::vector<uint32_t>> U(nf);
    P.height.assign(nf, 0);
    {
        std::vector<uint32_t> mark(ni, 0xffffffffu);
        for (uint32_t f = 0; f < nf; ++f) {
            std::vector<uint32_t> cand;
            auto add = [&](uint32_t p) {
                if (p >= send[f] && mark[p] != f) {
                    mark[p] = f;
                    cand.push_back(p);
                }
            };
            for (uint32_t v : D.S[f])
                for (uint32_t k = roff[v]; k < roff[v + 1]; ++k) add(D.pos[rcol[k]]);
            uint32_t h = 0;
            for (uint32_t c : D.kids[f]) {
                for (uint32_t p : U[c]) add(p);
                h = std::max(h, P.height[c] + 1);
            }
            P.height[f] = h;
            std::sort(cand.begin(), cand.end());
            U[f].swap(cand);
        }
    }
    lap("update-sets");
    // fronts
    P.fronts.resize(nf);
    uint64_t off = 0, upd = 0;
    uint32_t idxo = 0, nl = 0, invo = 0, fmo = 0;
    for (uint32_t f = 0; f < nf; ++f) {
        MfFront &F = P.fronts[f];
        F.s = (uint32_t)D.S[f].size();
        F.m = F.s + (uint32_t)U[f].size();
        if (F.m > kMfMaxFront) {
            dgx_set_error("coarse solver: a front of %u nodes exceeds %u (coarsest level too large)", F.m,
                          kMfMaxFront);
            throw DgxError();
        }
        F.off = off;
        off += (uint64_t)F.m * F.m;
        F.idx = idxo;
        idxo += F.m;
        F.upd = (uint32_t)upd;
        upd += F.m - F.s;
        F.ws = sstart[f];
        F.ch0 = nl;
        F.nch = (uint32_t)D.kids[f].size();
        nl += F.nch;
        for (uint32_t c : D.kids[f]) {
            invo += F.m;
            fmo += P.fronts[c].m - P.fronts[c].s;
        }
        P.max_m = std::max(P.max_m, F.m);
        P.max_s = std::max(P.max_s, F.s);
    }
    if (off > max_entries) {
        dgx_set_error("coarse solver: the frontal matrices need %llu doubles (limit %llu)", (unsigned long long)off,
                      (unsigned long long)max_entries);
        throw DgxError();
    }
    P.front_entries = off;
    P.upd_len = upd;
    P.idx.resize(idxo);
    P.links.resize(nl);
    P.inv.assign(invo, -1);
    P.fmap.resize(fmo);
    // local positions: pm[elimination position] = place in the current front
    std::vector<uint32_t> pm(ni, 0xffffffffu);
    P.a_dst.resize(nnz);
    invo = 0;
    fmo = 0;
    for (uint32_t f = 0; f < nf; ++f) {
        const MfFront &F = P.fronts[f];
        uint32_t *ix = P.idx.data() + F.idx;
        for (uint32_t p = 0; p < F.s; ++p) {
            ix[p] = P.inode[D.S[f][p]];
            pm[sstart[f] + p] = p;
        }
        for (uint32_t p = 0; p < (uint32_t)U[f].size(); ++p) {
            ix[F.s + p] = P.inode[at_pos[U[f][p]]];
            pm[U[f][p]] = F.s + p;
        }
        for (uint32_t l = 0; l < F.nch; ++l) {
            const uint32_t c = D.kids[f][l];
            MfLink &L = P.links[F.ch0 + l];
            L.child = c;
            L.inv = invo;
            L.fmap = fmo;
            const std::vector<uint32_t> &uc = U[c];
            for (uint32_t i = 0; i < (uint32_t)uc.size(); ++i) {
                const uint32_t p = pm[uc[i]];
                if (p >= F.m || (p < F.s ? sstart[f] + p : U[f][p - F.s]) != uc[i]) {
                    dgx_set_error("coarse solver: a child update index is missing from its parent front");
                    throw DgxError();
                }
                P.inv[invo + p] = (int32_t)i;
                P.fmap[fmo + i] = p;
            }
            invo += F.m;
            fmo += (uint32_t)uc.size();
        }
    }
    lap("fronts+maps");
    // The entries each front owns: (i, j) whose first-eliminated node is in S.
    // They lie in the rows of S (columns eliminated at S or later) and in the
    // rows of U (columns in S).  Fronts are independent: each range of fronts
    // keeps its own position map.
    std::atomic<bool> outside{false};
    parallel_ranges(nf, 16, [&](size_t lo, size_t hi, unsigned) {
        std::vector<uint32_t> lpm(ni, 0xffffffffu);
        for (size_t f = lo; f < hi; ++f) {
            const MfFront &F = P.fronts[f];
            for (uint32_t p = 0; p < F.s; ++p) lpm[sstart[f] + p] = p;
            for (uint32_t p = 0; p < (uint32_t)U[f].size(); ++p) lpm[U[f][p]] = F.s + p;
            auto put = [&](uint32_t k, uint32_t pi, uint32_t pj) {
                const uint32_t a = lpm[pi], b = lpm[pj];
                if (b >= F.m || (b < F.s ? sstart[f] + b : U[f][b - F.s]) != pj) {
                    outside = true;
                    return;
                }
                P.a_dst[k] = F.off + (uint64_t)a * F.m + b;
            };
            for (uint32_t v : D.S[f])
                for (uint32_t k = roff[v]; k < roff[v + 1]; ++k) {
                    const uint32_t pj = D.pos[rcol[k]];
                    if (pj >= sstart[f]) put(k, D.pos[v], pj);
                }
            for (uint32_t pu : U[f]) {
                const uint32_t u = at_pos[pu];
                for (uint32_t k = roff[u]; k < roff[u + 1]; ++k) {
                    const uint32_t pj = D.pos[rcol[k]];
                    if (pj >= sstart[f] && pj < send[f]) put(k, pu, pj);
                }
            }
        }
    });
    if (outside) {
        dgx_set_error("coarse solver: a matrix entry falls outside its front");
        throw DgxError();
    }
    lap("entry-places");
    // element-matrix contributions of each entry, in element order: row i's
    // entries come from its incident elements (in element order), each
    // element adding vals[e*64 + k*8 + j] to (ids[j^c], ids[k^c])
    P.a_off.assign(nnz + 1, 0);
    {
        std::vector<std::vector<uint32_t>> part(8);
        parallel_ranges(ni, 4096, [&](size_t lo, size_t hi, unsigned t) {
            std::vector<uint32_t> slot(ni, 0), cnt, start, &out = part[t];
            std::vector<std::pair<uint32_t, uint32_t>> buf;  // (entry in row, contribution)
            out.reserve((size_t)(ioff[hi] - ioff[lo]) * 8);
            for (uint32_t i = (uint32_t)lo; i < hi; ++i) {
                const uint32_t r0 = roff[i], rn = roff[i + 1] - r0;
                for (uint32_t k = r0; k < roff[i + 1]; ++k) slot[rcol[k]] = k - r0;
                cnt.assign(rn + 1, 0);
                buf.clear();
                for (uint32_t q = ioff[i]; q < ioff[i + 1]; ++q) {
                    const uint32_t e = inc[q] >> 3, jj = inc[q] & 7u;
                    const uint32_t c = (meta[e] >> 8) & 7u, j = jj ^ c;  // row i is child-frame vertex j
                    const int32_t *ids = &eid[8 * (size_t)e];
                    for (uint32_t kk = 0; kk < 8; ++kk) {
                        const int32_t col = ids[kk];
                        if (col < 0) continue;
                        const uint32_t k = kk ^ c, a = slot[col];
                        buf.push_back({a, (uint32_t)(e * 64 + k * 8 + j)});
                        ++cnt[a + 1];
                    }
                }
                for (uint32_t a = 0; a < rn; ++a) cnt[a + 1] += cnt[a];
                const size_t base = out.size();
                out.resize(base + buf.size());
                start.assign(cnt.begin(), cnt.end());
                for (const auto &pr : buf) out[base + start[pr.first]++] = pr.second;  // stable: element order
                for (uint32_t a = 0; a < rn; ++a) P.a_off[r0 + a + 1] = cnt[a + 1] - cnt[a];
            }
        });
        for (uint32_t k = 0; k < nnz; ++k) P.a_off[k + 1] += P.a_off[k];
        P.a_src.clear();
        P.a_src.reserve(P.a_off[nnz]);
        for (auto &pv : part) P.a_src.insert(P.a_src.end(), pv.begin(), pv.end());
    }
    lap("contributions");
    // levels by height, with their work items and scratch
    uint32_t H = 0;
    for (uint32_t f = 0; f < nf; ++f) H = std::max(H, P.height[f] + 1);
    P.lvl_fronts.assign(H, {});
    P.lvl_fwd.assign(H, {});
    P.lvl_bwd.assign(H, {});
    P.lvl_gather.assign(H, {});
    P.lvl_max_m.assign(H, 0);
    P.lvl_max_s.assign(H, 0);
    P.lvl_tc.assign(H, 0);
    for (uint32_t f = 0; f < nf; ++f) {
        const uint32_t h = P.height[f];
        MfFront &F = P.fronts[f];
        F.dinv = (uint32_t)P.lvl_fronts[h].size();
        F.tc = P.lvl_tc[h];
        P.lvl_tc[h] += 64ull * F.m;
        P.lvl_fronts[h].push_back(f);
        P.lvl_max_m[h] = std::max(P.lvl_max_m[h], F.m);
        P.lvl_max_s[h] = std::max(P.lvl_max_s[h], F.s);
        const uint32_t u = F.m - F.s;
        if (u == 0) P.lvl_fwd[h].push_back({f, 0});
        for (uint32_t r = 0; r < u; r += 16) P.lvl_fwd[h].push_back({f, r});
        for (uint32_t r = 0; r < F.s; r += 16) P.lvl_bwd[h].push_back({f, r});
        if (F.nch)
            for (uint32_t r = 0; r < F.m; r += 4) P.lvl_gather[h].push_back({f, r});
    }
    for (uint32_t h = 0; h < H; ++h) {
        P.tc_max = std::max(P.tc_max, P.lvl_tc[h]);
        P.dinv_max = std::max(P.dinv_max, (uint32_t)P.lvl_fronts[h].size());
    }
    if (P.dinv_max > 65535) {  // fronts of a level index grid dimension z of the update launch
        dgx_set_error("coarse solver: %u fronts in one tree level (limit 65535)", P.dinv_max);
        throw DgxError();
    }
    lap("levels");
    if (timing) fprintf(stderr, "mf_plan: %u interior nodes, %zu elements, %u fronts, ms:%s\n", ni, E, nf, tlog.c_str());
    return P;
}

// ------------------------------------------------------------- kernels ----
namespace {

constexpr int MB = 256;
constexpr uint32_t kGJ = 32;                       // pivots per block step
constexpr uint32_t kGJD = 1024, kGJR = kGJD / kGJ;  // pivot-block inverse: one entry per thread
constexpr uint32_t kGPC = 32, kGPG = MB / kGPC;     // panel columns per workgroup
constexpr uint32_t kGJT = 64;                       // update tile

__global__ void k_mf_fill0(double *F, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * MB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * MB) F[i] = 0.0;
}

// Each matrix entry sums its element-matrix contributions in element order and
// is stored at its place in its front (the places are distinct).
__global__ void k_mf_assemble(const uint32_t *__restrict__ a_off, const uint32_t *__restrict__ a_src,
                              const uint64_t *__restrict__ a_dst, uint32_t nnz, const double *__restrict__ vals,
                              double *F) {
    const uint32_t k = blockIdx.x * MB + threadIdx.x;
    if (k >= nnz) return;
    double v = 0.0;
    for (uint32_t i = a_off[k]; i < a_off[k + 1]; ++i) v += vals[a_src[i]];
    F[a_dst[k]] = v;
}

// Extend-add: each parent row adds its children's update rows, child by child.
// Four rows per workgroup, one wave per row.
__global__ void __launch_bounds__(MB) k_mf_gather(const MfItem *__restrict__ items, const MfFront *__restrict__ fr,
                                                  const MfLink *__restrict__ links, const int32_t *__restrict__ inv,
                                                  const uint32_t *__restrict__ fmap, double *F) {
    const MfItem it = items[blockIdx.x];
    const MfFront f = fr[it.front];
    const uint32_t p = it.r0 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (p >= f.m) return;
    double *Fr = F + f.off + (uint64_t)p * f.m;
    for (uint32_t l = 0; l < f.nch; ++l) {
        const MfLink L = links[f.ch0 + l];
        const int32_t ip = inv[L.inv + p];
        if (ip < 0) continue;
        const MfFront c = fr[L.child];
        const uint32_t uc = c.m - c.s;
        const double *Ur = F + c.off + (uint64_t)(c.s + ip) * c.m + c.s;
        for (uint32_t q = lane; q < uc; q += 64) Fr[fmap[L.fmap + q]] += Ur[q];
    }
}

// Inverse of the pivot block (k0, k0) of every front of the level whose
// separator reaches past k0, by Gauss-Jordan in LDS (rows/columns past the
// separator are an identity).  *bad is set for a pivot that is not positive.
__global__ void __launch_bounds__(kGJD) k_mf_gj_diag(const double *F, const MfFront *__restrict__ fr,
                                                     const uint32_t *__restrict__ lvl, uint32_t k0, double *dinv,
                                                     uint32_t *bad) {
    const MfFront f = fr[lvl[blockIdx.x]];
    if (k0 >= f.s) return;
    __shared__ double D[kGJ][kGJ + 1];
    const uint32_t kb = min(kGJ, f.s - k0), c = threadIdx.x % kGJ, r = threadIdx.x / kGJ;
    D[r][c] = (r < kb && c < kb) ? F[f.off + (uint64_t)(k0 + r) * f.m + k0 + c] : (r == c ? 1.0 : 0.0);
    for (uint32_t p = 0; p < kb; ++p) {
        __syncthreads();
        const double piv = D[p][p], dpc = D[p][c], fr_ = D[r][p];
        __syncthreads();
        if (threadIdx.x == 0 && !(piv > 0.0 && piv < INFINITY)) atomicOr(bad, 1u);
        double v;
        if (r == p) v = (c == p) ? 1.0 / piv : dpc / piv;
        else v = (c == p) ? -fr_ / piv : D[r][c] - fr_ * (dpc / piv);
        D[r][c] = v;
    }
    __syncthreads();
    dinv[(size_t)f.dinv * (kGJ * kGJ) + r * kGJ + c] = D[r][c];
}

// T[q][j] = sum_c Dinv[q][c] F[k0+c][j] for every column j, and C[i][c] =
// F[i][k0+c] for every row i, before the update overwrites them.
__global__ void __launch_bounds__(MB) k_mf_gj_panels(const double *F, const MfFront *__restrict__ fr,
                                                     const uint32_t *__restrict__ lvl, uint32_t k0,
                                                     const double *dinv, double *TC) {
    const MfFront f = fr[lvl[blockIdx.y]];
    const uint32_t n = f.m;
    if (k0 >= f.s || blockIdx.x * kGPC >= n) return;
    const uint32_t kb = min(kGJ, f.s - k0);
    const double *A = F + f.off;
    double *T = TC + f.tc, *C = T + (size_t)kGJ * n;
    __shared__ double Ds[kGJ * kGJ];
    __shared__ double As[kGJ][kGPC + 1];
    for (uint32_t t = threadIdx.x; t < kGJ * kGJ; t += MB) Ds[t] = dinv[(size_t)f.dinv * (kGJ * kGJ) + t];
    const uint32_t jj = threadIdx.x % kGPC, g = threadIdx.x / kGPC;
    const uint32_t j = blockIdx.x * kGPC + jj;
    for (uint32_t c = g; c < kGJ; c += kGPG) As[c][jj] = (c < kb && j < n) ? A[(size_t)(k0 + c) * n + j] : 0.0;
    __syncthreads();
    if (j < n)
        for (uint32_t q = g; q < kb; q += kGPG) {
            double t = 0.0;
#pragma unroll 8
            for (uint32_t c = 0; c < kGJ; ++c) t += Ds[q * kGJ + c] * As[c][jj];
            T[(size_t)q * n + j] = t;
        }
    for (uint32_t t = threadIdx.x; t < kGPC * kGJ; t += MB) {
        const uint32_t i = blockIdx.x * kGPC + t / kGJ, c = t % kGJ;
        if (i < n && c < kb) C[(size_t)i * kGJ + c] = A[(size_t)i * n + k0 + c];
    }
}

// The rank-kb update of a 64 x 64 tile and the block's new rows, columns and
// pivot block: F[i,j] -= C[i,:] T[:,j] (i, j outside K), F[K,j] = T[:,j],
// F[i,K] = -C[i,:] Dinv, F[K,K] = Dinv.
__global__ void __launch_bounds__(MB) k_mf_gj_update(double *F, const MfFront *__restrict__ fr,
                                                     const uint32_t *__restrict__ lvl, uint32_t k0,
                                                     const double *dinv, const double *TC) {
    const MfFront f = fr[lvl[blockIdx.z]];
    const uint32_t n = f.m;
    const uint32_t i0 = blockIdx.y * kGJT, j0 = blockIdx.x * kGJT;
    if (k0 >= f.s || i0 >= n || j0 >= n) return;
    const uint32_t kb = min(kGJ, f.s - k0);
    double *A = F + f.off;
    const double *T = TC + f.tc, *C = T + (size_t)kGJ * n, *Dv = dinv + (size_t)f.dinv * (kGJ * kGJ);
    __shared__ double Cs[kGJT][kGJ + 1];
    __shared__ double Ts[kGJ][kGJT + 1];
    __shared__ double Ds[kGJ][kGJ + 1];
    for (uint32_t t = threadIdx.x; t < kGJT * kGJ; t += MB) {
        const uint32_t rr = t / kGJ, cc = t % kGJ, i = i0 + rr;
        Cs[rr][cc] = (i < n && cc < kb) ? C[(size_t)i * kGJ + cc] : 0.0;
        const uint32_t q = t / kGJT, jj = t % kGJT, j = j0 + jj;
        Ts[q][jj] = (j < n && q < kb) ? T[(size_t)q * n + j] : 0.0;
    }
    const bool holds_k = j0 < k0 + kb && k0 < j0 + kGJT;
    if (holds_k)
        for (uint32_t t = threadIdx.x; t < kGJ * kGJ; t += MB) Ds[t / kGJ][t % kGJ] = Dv[t];
    __syncthreads();
    const uint32_t tr = (threadIdx.x / 16) * 4, tc = (threadIdx.x % 16) * 4;
    double acc[4][4] = {};
#pragma unroll 8
    for (uint32_t c = 0; c < kGJ; ++c) {
        double cv[4], tv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            cv[u] = Cs[tr + u][c];
            tv[u] = Ts[c][tc + u];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[u][v] += cv[u] * tv[v];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const uint32_t i = i0 + tr + u;
        if (i >= n) continue;
        const bool iK = i >= k0 && i < k0 + kb;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const uint32_t j = j0 + tc + v;
            if (j >= n) continue;
            const bool jK = j >= k0 && j < k0 + kb;
            double *a = &A[(size_t)i * n + j];
            if (iK && jK) {
                *a = Dv[(i - k0) * kGJ + (j - k0)];
            } else if (iK) {
                *a = Ts[i - k0][tc + v];
            } else if (jK) {
                double s = 0.0;
                for (uint32_t c = 0; c < kb; ++c) s += Cs[tr + u][c] * Ds[c][j - k0];
                *a = -s;
            } else {
                *a -= acc[u][v];
            }
        }
    }
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Each wave of a 16-row work item takes 4 consecutive rows: acc[q] = row
// (rb + q) . v over columns [0, ncol), rows of stride ld starting at A.  A lane
// keeps 4 rows x 4 column chunks of loads in flight, so a row costs a few
// memory round trips instead of one per 64 columns; the order of the sums is
// fixed.
__device__ __forceinline__ void rows_dot4(const double *A, uint64_t ld, uint32_t rb, uint32_t r_end, uint32_t ncol,
                                          const double *vl, uint32_t lane, double acc[4]) {
    bool ok[4];
    const double *Ar[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        ok[q] = rb + q < r_end;
        Ar[q] = A + (uint64_t)(ok[q] ? rb + q : rb) * ld;
        acc[q] = 0.0;
    }
    for (uint32_t j0 = lane; j0 < ncol; j0 += 64 * 4) {
        double a[4][4], v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t j = j0 + 64 * u;
            const bool in = j < ncol;
            v[u] = in ? vl[j] : 0.0;
#pragma unroll
            for (int q = 0; q < 4; ++q) a[q][u] = (in && ok[q]) ? Ar[q][j] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int u = 0; u < 4; ++u) acc[q] = fma(a[q][u], v[u], acc[q]);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = wave_sum(acc[q]);
}

// Forward sweep of one level: w_S = b_S + children's updates (kept in ws for
// the backward sweep), then the update rows upd = w_U + P21 w_S, with
// P21 = -F21 F11^-1 in rows s.. of the front.  16 rows per workgroup.
template <typename T>
__global__ void __launch_bounds__(MB) k_mf_fwd(const MfItem *__restrict__ items, const MfFront *__restrict__ fr,
                                               const MfLink *__restrict__ links, const int32_t *__restrict__ inv,
                                               const uint32_t *__restrict__ idx, const double *__restrict__ F,
                                               const T *__restrict__ b, double *ws, double *upd) {
    extern __shared__ double wl[];
    const MfItem it = items[blockIdx.x];
    const MfFront f = fr[it.front];
    for (uint32_t p = threadIdx.x; p < f.s; p += MB) {
        double v = (double)b[idx[f.idx + p]];
        for (uint32_t l = 0; l < f.nch; ++l) {
            const MfLink L = links[f.ch0 + l];
            const int32_t ip = inv[L.inv + p];
            if (ip >= 0) v += upd[fr[L.child].upd + ip];
        }
        wl[p] = v;
        if (it.r0 == 0) ws[f.ws + p] = v;
    }
    __syncthreads();
    const uint32_t u = f.m - f.s, lane = threadIdx.x & 63;
    const uint32_t rb = it.r0 + (threadIdx.x >> 6) * 4, r1 = min(it.r0 + 16u, u);
    if (rb >= r1) return;
    double acc[4];
    rows_dot4(F + f.off + (uint64_t)f.s * f.m, f.m, rb, r1, f.s, wl, lane, acc);
    if (lane < 4 && rb + lane < r1) {
        const uint32_t r = rb + lane;
        double wu = 0.0;
        for (uint32_t l = 0; l < f.nch; ++l) {
            const MfLink L = links[f.ch0 + l];
            const int32_t ip = inv[L.inv + f.s + r];
            if (ip >= 0) wu += upd[fr[L.child].upd + ip];
        }
        const double a = lane == 0 ? acc[0] : lane == 1 ? acc[1] : lane == 2 ? acc[2] : acc[3];
        upd[f.upd + r] = wu + a;
    }
}

// Backward sweep of one level: x_S = P11 w_S - P12 x_U (rows 0..s-1 of the
// front hold [F11^-1, F11^-1 F12]); x_U are ancestors' nodes, solved already.
__global__ void __launch_bounds__(MB) k_mf_bwd(const MfItem *__restrict__ items, const MfFront *__restrict__ fr,
                                               const uint32_t *__restrict__ idx, const double *__restrict__ F,
                                               const double *__restrict__ ws, double *x) {
    extern __shared__ double vl[];
    const MfItem it = items[blockIdx.x];
    const MfFront f = fr[it.front];
    for (uint32_t j = threadIdx.x; j < f.m; j += MB) vl[j] = j < f.s ? ws[f.ws + j] : -x[idx[f.idx + j]];
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t rb = it.r0 + (threadIdx.x >> 6) * 4, r1 = min(it.r0 + 16u, f.s);
    if (rb >= r1) return;
    double acc[4];
    rows_dot4(F + f.off, f.m, rb, r1, f.m, vl, lane, acc);
    if (lane < 4 && rb + lane < r1) {
        const double a = lane == 0 ? acc[0] : lane == 1 ? acc[1] : lane == 2 ? acc[2] : acc[3];
        x[idx[f.idx + rb + lane]] = a;
    }
}

template <typename T>
__global__ void k_mf_finish(const uint8_t *__restrict__ bdy, const double *__restrict__ x, const T *__restrict__ b,
                            uint32_t n, T *z) {
    const uint32_t i = blockIdx.x * MB + threadIdx.x;
    if (i < n) z[i] = bdy[i] ? b[i] : (T)x[i];
}

template <typename T>
void upload(DevBuf<T> &d, const std::vector<T> &h, hipStream_t s) {
    d.alloc(h.size(), s);
    if (!h.empty()) HIP_OK(hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
}
template <typename T>
void download(std::vector<T> &h, const T *d, size_t n, hipStream_t s) {
    h.resize(n);
    if (n) HIP_OK(hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, s));
}

}  // namespace

CoarseInputs coarse_download(const DaDevice &d, hipStream_t s) {
    CoarseInputs in;
    in.n = d.n_nodes;
    in.E = d.n_elem;
    std::vector<uint32_t> n2l, x, y, z;
    download(in.conn, d.conn.p, 8 * in.E, s);
    download(in.meta, d.meta.p, in.E, s);
    download(in.bdy, d.bdy_node.p, in.n, s);
    download(n2l, d.n2l.p, in.n, s);
    download(x, d.x.p, d.n_local, s);
    download(y, d.y.p, d.n_local, s);
    download(z, d.z.p, d.n_local, s);
    HIP_OK(hipStreamSynchronize(s));
    in.xyz.resize(3 * (size_t)in.n);
    for (uint32_t i = 0; i < in.n; ++i) {
        in.xyz[3 * (size_t)i] = x[n2l[i]];
        in.xyz[3 * (size_t)i + 1] = y[n2l[i]];
        in.xyz[3 * (size_t)i + 2] = z[n2l[i]];
    }
    return in;
}

MfPlan coarse_plan(const CoarseInputs &in, uint32_t dense) {
    // the fronts may take at most 4 GiB of doubles
    return mf_plan(in.n, in.E, in.conn.data(), in.meta.data(), in.bdy.data(), in.xyz.data(), (4ull << 30) / 8, 128,
                   dense);
}

std::unique_ptr<CoarseSolver> coarse_upload(MfPlan &&plan, uint32_t n, size_t E, hipStream_t s) {
    std::unique_ptr<CoarseSolver> cs(new CoarseSolver());
    cs->plan = std::move(plan);
    const MfPlan &P = cs->plan;
    cs->n = n;
    cs->ni = P.ni;
    cs->n_elem = E;
    cs->nfronts = (uint32_t)P.fronts.size();
    cs->levels = (uint32_t)P.lvl_fronts.size();
    upload(cs->fronts, P.fronts, s);
    upload(cs->links, P.links, s);
    upload(cs->inv, P.inv, s);
    upload(cs->fmap, P.fmap, s);
    upload(cs->idx, P.idx, s);
    upload(cs->a_off, P.a_off, s);
    upload(cs->a_src, P.a_src, s);
    upload(cs->a_dst, P.a_dst, s);
    cs->lvl_fronts.resize(cs->levels);
    cs->lvl_fwd.resize(cs->levels);
    cs->lvl_bwd.resize(cs->levels);
    cs->lvl_gather.resize(cs->levels);
    for (uint32_t h = 0; h < cs->levels; ++h) {
        upload(cs->lvl_fronts[h], P.lvl_fronts[h], s);
        upload(cs->lvl_fwd[h], P.lvl_fwd[h], s);
        upload(cs->lvl_bwd[h], P.lvl_bwd[h], s);
        upload(cs->lvl_gather[h], P.lvl_gather[h], s);
    }
    cs->F.alloc(P.front_entries, s);
    cs->upd.alloc(P.upd_len, s);
    cs->ws.alloc(P.ni, s);
    cs->x.alloc(n, s);
    HIP_OK(hipStreamSynchronize(s));  // the uploads read host memory of the plan
    return cs;
}

std::unique_ptr<CoarseSolver> coarse_analyse(const DaDevice &d, hipStream_t s, uint32_t dense) {
    const CoarseInputs in = coarse_download(d, s);
    return coarse_upload(coarse_plan(in, dense), in.n, in.E, s);
}

void coarse_factor(CoarseSolver &cs, const DaDevice &d, double lap_fac, double mass_fac, hipStream_t s) {
    const MfPlan &P = cs.plan;
    const uint32_t nnz = (uint32_t)P.a_dst.size();
    {
        DevBuf<double> vals((size_t)64 * cs.n_elem, s);
        if (cs.n_elem) da_element_matrices(d, lap_fac, mass_fac, vals.p, s);
        if (P.front_entries)
            k_mf_fill0<<<(unsigned)std::min<uint64_t>(grid_for(P.front_entries, MB), 8192), MB, 0, s>>>(
                cs.F.p, P.front_entries);
        if (nnz)
            k_mf_assemble<<<grid_for(nnz, MB), MB, 0, s>>>(cs.a_off.p, cs.a_src.p, cs.a_dst.p, nnz, vals.p,
                                                           cs.F.p);
        HIP_OK(hipGetLastError());
    }
    DevBuf<double> TC(std::max<uint64_t>(P.tc_max, 1), s), dinv((size_t)std::max(P.dinv_max, 1u) * kGJ * kGJ, s);
    DevBuf<uint32_t> bad(1, s);
    HIP_OK(hipMemsetAsync(bad.p, 0, 4, s));
    for (uint32_t h = 0; h < cs.levels; ++h) {
        const uint32_t nf = (uint32_t)P.lvl_fronts[h].size();
        if (!P.lvl_gather[h].empty())
            k_mf_gather<<<(unsigned)P.lvl_gather[h].size(), MB, 0, s>>>(cs.lvl_gather[h].p, cs.fronts.p,
                                                                        cs.links.p, cs.inv.p, cs.fmap.p, cs.F.p);
        const uint32_t mm = P.lvl_max_m[h];
        const dim3 pg(grid_for(mm, kGPC), nf), ug(grid_for(mm, kGJT), grid_for(mm, kGJT), nf);
        for (uint32_t k0 = 0; k0 < P.lvl_max_s[h]; k0 += kGJ) {
            k_mf_gj_diag<<<nf, kGJD, 0, s>>>(cs.F.p, cs.fronts.p, cs.lvl_fronts[h].p, k0, dinv.p, bad.p);
            k_mf_gj_panels<<<pg, MB, 0, s>>>(cs.F.p, cs.fronts.p, cs.lvl_fronts[h].p, k0, dinv.p, TC.p);
            k_mf_gj_update<<<ug, MB, 0, s>>>(cs.F.p, cs.fronts.p, cs.lvl_fronts[h].p, k0, dinv.p, TC.p);
        }
        HIP_OK(hipGetLastError());
    }
    uint32_t hbad = 0;
    dgx_read_back(&hbad, bad.p, 4, s);
    if (hbad) {
        dgx_set_error("multigrid: coarse matrix is not positive definite (a pivot is not positive)");
        throw DgxError();
    }
}

template <typename T>
void coarse_solve(CoarseSolver &cs, const uint8_t *bdy, const T *b, T *z, hipStream_t s) {
    const MfPlan &P = cs.plan;
    for (uint32_t h = 0; h < cs.levels; ++h)
        k_mf_fwd<T><<<(unsigned)P.lvl_fwd[h].size(), MB, P.lvl_max_s[h] * sizeof(double), s>>>(
            cs.lvl_fwd[h].p, cs.fronts.p, cs.links.p, cs.inv.p, cs.idx.p, cs.F.p, b, cs.ws.p, cs.upd.p);
    for (uint32_t h = cs.levels; h-- > 0;)
        k_mf_bwd<<<(unsigned)P.lvl_bwd[h].size(), MB, P.lvl_max_m[h] * sizeof(double), s>>>(
            cs.lvl_bwd[h].p, cs.fronts.p, cs.idx.p, cs.F.p, cs.ws.p, cs.x.p);
    if (cs.n) k_mf_finish<T><<<grid_for(cs.n, MB), MB, 0, s>>>(bdy, cs.x.p, b, cs.n, z);
    HIP_OK(hipGetLastError());
}
template void coarse_solve<double>(CoarseSolver &, const uint8_t *, const double *, double *, hipStream_t);
template void coarse_solve<float>(CoarseSolver &, const uint8_t *, const float *, float *, hipStream_t);

// ------------------------------------------------------- C ABI (host plan) ----
struct dgx_direct_plan_s {
    MfPlan plan;
};

extern "C" {

int dgx_direct_plan_host(uint32_t n, uint64_t n_elem, const uint32_t *conn, const uint16_t *meta, const uint8_t *bdy,
                         const uint32_t *xyz, uint32_t leaf, uint32_t dense, dgx_direct_plan *out,
                         dgx_direct_sizes *sz) {
    if (!out || !sz || (n_elem && (!conn || !meta)) || (n && (!bdy || !xyz))) {
        dgx_set_error("direct plan: null argument");
        return -1;
    }
    try {
        std::unique_ptr<dgx_direct_plan_s> h(new dgx_direct_plan_s());
        h->plan = mf_plan(n, n_elem, conn, meta, bdy, xyz, ~0ull, leaf ? leaf : 128, dense);
        const MfPlan &P = h->plan;
        sz->interior = P.ni;
        sz->fronts = P.fronts.size();
        sz->links = P.links.size();
        sz->levels = P.lvl_fronts.size();
        sz->entries = P.front_entries;
        sz->idx_len = P.idx.size();
        sz->inv_len = P.inv.size();
        sz->fmap_len = P.fmap.size();
        sz->a_nnz = P.a_dst.size();
        sz->a_src_len = P.a_src.size();
        sz->upd_len = P.upd_len;
        sz->max_m = P.max_m;
        sz->max_s = P.max_s;
        *out = h.release();
        return 0;
    } catch (DgxError &) {
        return -1;
    } catch (std::exception &e) {
        dgx_set_error("direct plan: %s", e.what());
        return -1;
    }
}

int dgx_direct_plan_export(dgx_direct_plan h, uint64_t *front9, uint64_t *link3, int32_t *inv, uint32_t *fmap,
                           uint32_t *idx, uint32_t *a_off, uint32_t *a_src, uint64_t *a_dst, uint32_t *inode) {
    if (!h) {
        dgx_set_error("direct plan: null handle");
        return -1;
    }
    const MfPlan &P = h->plan;
    if (front9)
        for (size_t f = 0; f < P.fronts.size(); ++f) {
            const MfFront &F = P.fronts[f];
            const uint64_t v[9] = {F.off, F.m, F.s, F.idx, F.upd, F.ws, F.ch0, F.nch, P.height[f]};
            std::memcpy(front9 + 9 * f, v, sizeof(v));
        }
    if (link3)
        for (size_t l = 0; l < P.links.size(); ++l) {
            link3[3 * l] = P.links[l].child;
            link3[3 * l + 1] = P.links[l].inv;
            link3[3 * l + 2] = P.links[l].fmap;
        }
    if (inv) std::copy(P.inv.begin(), P.inv.end(), inv);
    if (fmap) std::copy(P.fmap.begin(), P.fmap.end(), fmap);
    if (idx) std::copy(P.idx.begin(), P.idx.end(), idx);
    if (a_off) std::copy(P.a_off.begin(), P.a_off.end(), a_off);
    if (a_src) std::copy(P.a_src.begin(), P.a_src.end(), a_src);
    if (a_dst) std::copy(P.a_dst.begin(), P.a_dst.end(), a_dst);
    if (inode) std::copy(P.inode.begin(), P.inode.end(), inode);
    return 0;
}

int dgx_direct_plan_free(dgx_direct_plan h) {
    delete h;
    return 0;
}

}  // extern "C"
