// Host orchestration of the smfgpu numeric phase: plan creation (the
// per-analysis set-up of multifrontal_factor_serial,
// solver/src/multifrontal_serial.cpp:1105-1320 of the reference), the
// level-batched factorization, solves, refinement and the FactorKeep export.
#include "smfgpu_plan.hpp"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>

namespace smfgpu {

namespace {
thread_local std::string g_err;
}

int set_error(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

const char *last_error() { return g_err.c_str(); }

// device precision code of a SMFGPU_PREC_* request (DevPtrs::prec), -1
// when the request cannot be served
int prec_code(int32_t precision) {
    switch (precision) {
    case SMFGPU_PREC_FP64: return 0;
    case SMFGPU_PREC_FP32: return 1;
    case SMFGPU_PREC_BF16: return 2;
    case SMFGPU_PREC_BF16X3: return 3;
    default: return -1;
    }
}


}  // namespace smfgpu

using namespace smfgpu;

namespace {

DevPtrs dev_ptrs(smfgpu_plan_s *P) {
    DevPtrs d{};
    d.fronts = P->d_fronts.p;
    d.rows = P->d_rows.p;
    d.cmap = P->d_cmap.p;
    d.cmofs = P->d_cmofs.p;
    d.pcol_ptr = P->d_pcol_ptr.p;
    d.amap = P->d_amap.p;
    d.av = P->d_av.p;
    d.L = P->d_L.p;
    d.L32 = P->d_L32.p;
    d.L16h = P->d_L16h.p;
    d.L16l = P->d_L16l.p;
    d.CB = P->d_CB.p;
    d.flag = P->d_flag.p;
    d.prec = P->d_prec.p;
    d.junk = P->d_junk.p;
    d.colstat = nullptr;
    return d;
}

static void release_graph(smfgpu_plan_s *P) {
    for (int g = 0; g < 3; ++g) {
        if (P->fexec[g]) (void)hipGraphExecDestroy(P->fexec[g]);
        if (P->fgraph[g]) (void)hipGraphDestroy(P->fgraph[g]);
        P->fexec[g] = nullptr;
        P->fgraph[g] = nullptr;
    }
}

// Solve schedule (static for SPD). Per front, pivot blocks of
// kSolveBlock columns; step j of the forward sweep solves block j's triangle
// (one workgroup, after applying block j-1 to its rows) while 64-row chunks
// apply block j-1 to the rows below. The backward sweep runs the blocks in
// reverse: the rows below the pivots first (segments of kSolveSeg rows),
// then per block the triangle and 64-column chunks of the columns before it.
// A front whose last block has at most solve_fuse_rows rows below it lets
// the triangle's workgroup apply that block itself. Steps of all fronts of
// a level are merged (triangle items first: they are each step's critical
// path).
int solve_lists(smfgpu_plan_s *P) {
    const int32_t nlev = static_cast<int32_t>(P->level_fronts.size());
    const int32_t B = kSolveBlock, fuse = P->solve_fuse_rows;
    std::vector<SolveItem> fwd_items, bwd_items;
    P->fwd_steps.clear();
    P->bwd_steps.clear();
    P->fwd_lv.assign(nlev + 1, 0);
    P->bwd_lv.assign(nlev + 1, 0);
    std::vector<std::vector<SolveItem>> ftri, fchunk, btri, bchunk;
    auto put = [](std::vector<std::vector<SolveItem>> &v, size_t t, const SolveItem &it) {
        if (v.size() <= t) v.resize(t + 1);
        v[t].push_back(it);
    };
    auto flush = [](std::vector<std::vector<SolveItem>> &tri, std::vector<std::vector<SolveItem>> &chunk,
                    std::vector<SolveItem> &items, std::vector<std::pair<int64_t, int64_t>> &steps) {
        const size_t nt = std::max(tri.size(), chunk.size());
        tri.resize(nt);
        chunk.resize(nt);
        for (size_t t = 0; t < nt; ++t) {
            const int64_t first = static_cast<int64_t>(items.size());
            items.insert(items.end(), tri[t].begin(), tri[t].end());
            items.insert(items.end(), chunk[t].begin(), chunk[t].end());
            const int64_t cnt = static_cast<int64_t>(items.size()) - first;
            if (cnt > 0) steps.emplace_back(first, cnt);
        }
        tri.clear();
        chunk.clear();
    };
    for (int32_t lv = 0; lv < nlev; ++lv) {
        P->fwd_lv[lv] = static_cast<int32_t>(P->fwd_steps.size());
        P->bwd_lv[lv] = static_cast<int32_t>(P->bwd_steps.size());
        for (int32_t s : P->level_fronts[lv]) {
            const FrontDev &F = P->fronts[s];
            const int32_t f = F.f, p = F.p, q = f - p;
            if (p <= 0) continue;
            const int32_t nb = (p + B - 1) / B;
            // forward
            size_t t = 0;
            bool tail_done = false;
            for (int32_t j = 0; j < nb; ++j, ++t) {
                const int32_t tk = j * B, tkb = std::min(B, p - tk);
                const int32_t k = j > 0 ? tk - B : 0, kb = j > 0 ? B : 0;
                const bool tail = j == nb - 1 && q > 0 && q <= fuse;
                tail_done = tail_done || tail;
                put(ftri, t, SolveItem{s, k, kb, tk, tk + tkb, tk, tkb, tail ? 2 : 0});
                if (j > 0)
                    for (int32_t r = tk + tkb; r < f; r += 64)
                        put(fchunk, t, SolveItem{s, k, kb, r, std::min(r + 64, f), 0, 0, 0});
            }
            if (q > 0 && !tail_done) {
                const int32_t k = (nb - 1) * B;
                for (int32_t r = p; r < f; r += 64)
                    put(fchunk, t, SolveItem{s, k, p - k, r, std::min(r + 64, f), 0, 0, 0});
            }
            // backward
            const int32_t kL = (nb - 1) * B, tkbL = p - kL;
            t = 0;
            if (q > 0 && q <= kSolveSeg) {
                for (int32_t c = 0; c < kL; c += 64)
                    put(bchunk, 0, SolveItem{s, p, q, c, std::min(c + 64, kL), 0, 0, 0});
                put(btri, 0, SolveItem{s, p, q, kL, p, kL, tkbL, 0});
                t = 1;
            } else if (q > 0) {
                for (int32_t r = p; r < f; r += kSolveSeg)
                    for (int32_t c = 0; c < p; c += 64)
                        put(bchunk, 0, SolveItem{s, r, std::min(kSolveSeg, f - r), c, std::min(c + 64, p), 0, 0, 1});
                put(btri, 1, SolveItem{s, 0, 0, 0, 0, kL, tkbL, 0});
                t = 2;
            } else {
                put(btri, 0, SolveItem{s, 0, 0, 0, 0, kL, tkbL, 0});
                t = 1;
            }
            for (int32_t j = nb - 2; j >= 0; --j, ++t) {
                const int32_t tk = j * B, k = tk + B, kb = std::min(B, p - k);
                put(btri, t, SolveItem{s, k, kb, tk, tk + B, tk, B, 0});
                for (int32_t c = 0; c < tk; c += 64)
                    put(bchunk, t, SolveItem{s, k, kb, c, std::min(c + 64, tk), 0, 0, 0});
            }
        }
        flush(ftri, fchunk, fwd_items, P->fwd_steps);
        flush(btri, bchunk, bwd_items, P->bwd_steps);
    }
    P->fwd_lv[nlev] = static_cast<int32_t>(P->fwd_steps.size());
    P->bwd_lv[nlev] = static_cast<int32_t>(P->bwd_steps.size());
    for (const auto &st : P->fwd_steps)
        if (st.second > INT32_MAX) return set_error(SMFGPU_ERR_ALLOC, "plan: solve step too large");
    for (const auto &st : P->bwd_steps)
        if (st.second > INT32_MAX) return set_error(SMFGPU_ERR_ALLOC, "plan: solve step too large");
    int rc;
    if ((rc = P->d_fwd_items.upload(fwd_items)) || (rc = P->d_bwd_items.upload(bwd_items))) return rc;
    // the persistent sweeps' sequences
    std::vector<int32_t> fstep(fwd_items.size()), fcnt, bstep, bcnt;
    for (size_t t = 0; t < P->fwd_steps.size(); ++t) {
        for (int64_t i = 0; i < P->fwd_steps[t].second; ++i) fstep[P->fwd_steps[t].first + i] = static_cast<int32_t>(t);
        fcnt.push_back(static_cast<int32_t>(P->fwd_steps[t].second));
    }
    std::vector<SolveItem> bseq;
    bseq.reserve(bwd_items.size());
    for (int32_t lv = nlev - 1; lv >= 0; --lv)
        for (int32_t i = P->bwd_lv[lv]; i < P->bwd_lv[lv + 1]; ++i) {
            const auto &stp = P->bwd_steps[i];
            for (int64_t e = 0; e < stp.second; ++e) {
                bseq.push_back(bwd_items[stp.first + e]);
                bstep.push_back(static_cast<int32_t>(bcnt.size()));
            }
            bcnt.push_back(static_cast<int32_t>(stp.second));
        }
    const size_t nctr = 1 + std::max(fcnt.size(), bcnt.size());
    if ((rc = P->d_bwd_seq.upload(bseq)) || (rc = P->d_fwd_step_of.upload(fstep)) ||
        (rc = P->d_bwd_step_of.upload(bstep)) || (rc = P->d_fwd_step_cnt.upload(fcnt)) ||
        (rc = P->d_bwd_step_cnt.upload(bcnt)) || (rc = P->d_sweep_ctr.alloc(nctr)))
        return rc;
    if (!P->d_sweep_err.p) {
        if ((rc = P->d_sweep_err.alloc(1)) || (rc = P->h_sweep_err.assign(1, 0))) return rc;
        HIP_OK(hipMemset(P->d_sweep_err.p, 0, sizeof(int32_t)));
    }
    return SMFGPU_OK;
}

// SPD launch sequence (static): per level, the assembly, then per outer
// panel of panel_outer pivot columns and per 64-column block of it: the
// diagonal Cholesky, the panel TRSM and the Schur update of the rest of the
// outer panel (K = 64); after the outer panel, the Schur update of the
// remaining pivot columns with the whole outer panel (K = panel_outer);
// after the last pivot, the contribution block's update in K segments of
// cb_seg columns. Items of all fronts of the level share each launch.
int dense_schedule(smfgpu_plan_s *P) {
    const int32_t nlev = static_cast<int32_t>(P->level_fronts.size());
    std::vector<PotrfItem> potrf;
    std::vector<TrsmItem> trsm;
    std::vector<SyrkItem> syrk;
    P->selfr.clear();
    // one launch sequence per outer panel width: FP64 factorizations take
    // panel_outer, reduced-precision ones panel_outer_red (their trailing
    // updates of the pivot columns accumulate K = outer products in FP32, so
    // a narrower panel keeps that sum shorter); both index the same item
    // arrays
    auto build = [&](const int32_t outer, std::vector<Launch> &launches, std::vector<int64_t> &sflops) {
    launches.clear();
    sflops.clear();
    // dense-kernel items and the launch sequence
    // tiles (it, jt) of the lower part of rows [jlo, f) x columns [jlo, jhi),
    // in super-blocks of kSyrkGroup x kSyrkGroup tiles: the kernel gives each
    // XCD a contiguous range of items, so the tiles resident on one XCD
    // share their row and column panels in its L2
    auto add_syrk = [&](int32_t s, int32_t jlo, int32_t jhi, int32_t t0, int32_t t1) {
        const FrontDev &F = P->fronts[s];
        const int32_t nc = (jhi - jlo + kTile - 1) / kTile, nr = (F.f - jlo + kTile - 1) / kTile;
        for (int32_t C = 0; C < nc; C += kSyrkGroup)
            for (int32_t R = C; R < nr; R += kSyrkGroup)
                for (int32_t cb = C; cb < std::min(C + kSyrkGroup, nc); ++cb)
                    for (int32_t rb = std::max(R, cb); rb < std::min(R + kSyrkGroup, nr); ++rb)
                        syrk.push_back(SyrkItem{s, jlo + rb * kTile, jlo + cb * kTile, jhi, t0, t1, 0, 0});
    };
    // a Schur launch: the items of the fronts that always run in FP64
    // first, then those of the fronts that may run below it (f >=
    // red_min), counted in Launch::red
    const int32_t red_min = P->sched_red_min;
    auto syrk_launch = [&](int32_t lv, const std::vector<int32_t> &fl, auto &&emit) {
        const int64_t y0 = static_cast<int64_t>(syrk.size());
        for (int32_t s : fl)
            if (P->fronts[s].f < red_min) emit(s);
        const int64_t yr = static_cast<int64_t>(syrk.size());
        for (int32_t s : fl)
            if (P->fronts[s].f >= red_min) emit(s);
        const int64_t y1 = static_cast<int64_t>(syrk.size());
        launches.push_back(Launch{K_SYRK, lv, y0, y1 - y0, y1 - yr});
    };
    for (int32_t lv = 0; lv < nlev; ++lv) {
        const auto &fl = P->level_fronts[lv];
        launches.push_back(Launch{K_ASSEMBLE, lv, P->asm_first[lv], P->asm_first[lv + 1] - P->asm_first[lv], 0});
        // the selector's candidates of the level (fronts that may run below
        // FP64), between their assembly and their factorization
        const int64_t s0 = static_cast<int64_t>(P->selfr.size());
        for (int32_t s : fl)
            if (P->fronts[s].f >= red_min)
                P->selfr.push_back(SelFront{s, P->fronts[s].f, P->fronts[s].p, 0, P->fronts[s].col0});
        launches.push_back(Launch{K_SELECT, lv, s0, static_cast<int64_t>(P->selfr.size()) - s0, 0});
        int32_t maxp = 0;
        for (int32_t s : fl) maxp = std::max(maxp, P->fronts[s].p);
        for (int32_t K = 0; K < maxp; K += outer) {
            for (int32_t k = K; k < std::min(K + outer, maxp); k += kPanelInner) {
                const int64_t p0 = static_cast<int64_t>(potrf.size());
                const int64_t t0 = static_cast<int64_t>(trsm.size());
                for (int32_t s : fl) {
                    const FrontDev &F = P->fronts[s];
                    if (F.p <= k) continue;
                    const int32_t kb = std::min(kPanelInner, F.p - k);
                    potrf.push_back(PotrfItem{s, k, kb, 0});
                    for (int32_t r0 = k + kb; r0 < F.f; r0 += kPanelInner) trsm.push_back(TrsmItem{s, k, kb, r0});
                }
                launches.push_back(Launch{K_POTRF, lv, p0, static_cast<int64_t>(potrf.size()) - p0, 0});
                launches.push_back(Launch{K_TRSM, lv, t0, static_cast<int64_t>(trsm.size()) - t0, 0});
                syrk_launch(lv, fl, [&](int32_t s) {
                    const FrontDev &F = P->fronts[s];
                    if (F.p <= k) return;
                    const int32_t kb = std::min(kPanelInner, F.p - k);
                    const int32_t kend = std::min(K + outer, F.p);
                    if (k + kb < kend) add_syrk(s, k + kb, kend, k, k + kb);
                });
            }
            // trailing update of the remaining pivot columns only (all
            // their rows, the ones below the pivots included)
            syrk_launch(lv, fl, [&](int32_t s) {
                const FrontDev &F = P->fronts[s];
                if (F.p <= K) return;
                const int32_t kend = std::min(K + outer, F.p);
                if (kend < F.p) add_syrk(s, kend, F.p, K, kend);
            });
        }
        // the contribution blocks, once per cb_seg pivot columns (the
        // reference's final schur_update, ldlt_kernel.cpp:1135): far less
        // read-modify-write traffic than one update per outer panel
        for (int32_t t0 = 0; t0 < maxp; t0 += P->cb_seg) {
            syrk_launch(lv, fl, [&](int32_t s) {
                const FrontDev &F = P->fronts[s];
                if (F.p <= t0 || F.q <= 0) return;
                add_syrk(s, F.p, F.f, t0, std::min(t0 + P->cb_seg, F.p));
            });
        }
    }
    launches.erase(std::remove_if(launches.begin(), launches.end(), [](const Launch &l) { return l.count <= 0; }),
                   launches.end());
    // syrk flops per launch (algorithmic: 2 * entries updated * panel width)
    for (const Launch &l : launches) {
        if (l.kind != K_SYRK) {
            sflops.push_back(0);
            continue;
        }
        int64_t fl = 0;
        for (int64_t i = l.first; i < l.first + l.count; ++i) {
            const SyrkItem &y = syrk[i];
            const FrontDev &F = P->fronts[y.s];
            const int64_t jh = std::min<int64_t>(y.j0 + kTile, std::min(y.j1, F.f));
            const int64_t ih = std::min<int64_t>(y.i0 + kTile, F.f);
            int64_t cnt = 0;
            for (int64_t j = y.j0; j < jh; ++j) cnt += std::max<int64_t>(0, ih - std::max<int64_t>(j, y.i0));
            fl += 2 * cnt * (y.t1 - y.t0);
        }
        sflops.push_back(fl);
    }
    };
    build(P->panel_outer, P->launches, P->syrk_flops_of_launch);
    if (P->panel_outer_red != P->panel_outer) {
        build(P->panel_outer_red, P->launches_red, P->syrk_flops_of_launch_red);
    } else {
        P->launches_red = P->launches;
        P->syrk_flops_of_launch_red = P->syrk_flops_of_launch;
    }
    int rc;
    if ((rc = P->d_potrf.upload(potrf)) || (rc = P->d_trsm.upload(trsm)) || (rc = P->d_syrk.upload(syrk)) ||
        (rc = P->d_selfr.upload(P->selfr)))
        return rc;
    release_graph(P);
    return SMFGPU_OK;
}

}  // namespace

SelPrm smfgpu::sel_prm(smfgpu_plan_s *P, const smfgpu_factor_options &opt) {
    SelPrm sp{};
    const bool autosel = opt.precision == SMFGPU_PREC_AUTO;
    sp.w = autosel ? P->d_selw.p : nullptr;
    sp.log_tol = static_cast<float>(std::log10(autosel ? opt.selector_tol : 1.0));
    int32_t mask = 0;  // device codes: bit 1 FP32, bit 2 BF16, bit 3 split BF16
    const int32_t am = opt.selector_mask != 0
                           ? opt.selector_mask
                           : (1 << SMFGPU_PREC_FP32) | (1 << SMFGPU_PREC_BF16) | (1 << SMFGPU_PREC_BF16X3);
    if (am & (1 << SMFGPU_PREC_FP32)) mask |= 2;
    if (am & (1 << SMFGPU_PREC_BF16)) mask |= 4;
    if (am & (1 << SMFGPU_PREC_BF16X3)) mask |= 8;
    sp.mask = mask;
    sp.scaled = opt.scaled != 0;
    sp.maxabs_bits = reinterpret_cast<const unsigned long long *>(P->d_red.p);
    return sp;
}

namespace {

int plan_build(smfgpu_plan_s *P, const smfgpu_analysis_desc *D) {
    const int32_t n = D->n, ns = D->nsup;
    P->n = n;
    P->nsup = ns;
    P->perm.assign(D->perm, D->perm + n);
    P->nnz_perm = D->pcol_ptr[n];
    P->n_orig = D->n_orig;
    std::vector<int32_t> col2sn(n, -1);
    P->fronts.resize(ns);
    P->front_flops.assign(ns, 0.0);
    int64_t lofs = 0, cbofs = 0, l16ofs = 0;
    for (int32_t s = 0; s < ns; ++s) {
        FrontDev &F = P->fronts[s];
        F.f = D->front_row_ptr[s + 1] - D->front_row_ptr[s];
        F.p = D->sn_end[s] - D->sn_start[s];
        F.q = F.f - F.p;
        F.col0 = D->sn_start[s];
        F.lofs = lofs;
        F.cbofs = cbofs;
        F.rofs = D->front_row_ptr[s];
        F.l16ofs = l16ofs;
        if (F.p <= 0 || F.q < 0) return set_error(SMFGPU_ERR_INVALID_ARG, "plan: bad supernode sizes");
        for (int32_t j = D->sn_start[s]; j < D->sn_end[s]; ++j) col2sn[j] = s;
        lofs += static_cast<int64_t>(F.f) * F.p;
        l16ofs += static_cast<int64_t>(F.f) * ((F.p + 7) / 8 * 8);
        cbofs += static_cast<int64_t>(F.q) * F.q;
        double fl = 0.0;
        for (int c = 0; c < F.p; ++c) {
            const double rem = static_cast<double>(F.f - c - 1);
            fl += rem + rem * (rem + 1.0);
        }
        P->front_flops[s] = fl;
        P->flops += fl;
    }
    P->factor_entries = lofs;
    P->l16_entries = l16ofs;
    P->cb_entries = cbofs;
    P->rows_total = D->front_row_ptr[ns];
    // levels (children before parents; supernodes are in postorder)
    P->level.assign(ns, 0);
    int32_t nlev = 0;
    for (int32_t s = 0; s < ns; ++s) {
        const int32_t par = D->sn_parent[s];
        if (par >= 0) {
            if (par <= s) return set_error(SMFGPU_ERR_INVALID_ARG, "plan: supernodes not in postorder");
            P->level[par] = std::max(P->level[par], P->level[s] + 1);
        }
        nlev = std::max(nlev, P->level[s] + 1);
    }
    P->level_fronts.assign(nlev, {});
    for (int32_t s = 0; s < ns; ++s) P->level_fronts[P->level[s]].push_back(s);
    // local positions: amap for the permuted entries, cmap for each child
    std::vector<int32_t> pos(n, -1);
    std::vector<int32_t> amap(P->nnz_perm, 0);
    std::vector<int64_t> cmofs(ns + 1, 0);
    for (int32_t s = 0; s < ns; ++s) cmofs[s + 1] = cmofs[s] + P->fronts[s].q;
    std::vector<int32_t> cmap(cmofs[ns]);
    for (int32_t s = 0; s < ns; ++s) {
        const FrontDev &F = P->fronts[s];
        const int32_t *fr = D->front_row_idx + F.rofs;
        for (int t = 0; t < F.f; ++t) pos[fr[t]] = t;
        for (int t = 0; t < F.p; ++t)
            if (fr[t] != F.col0 + t) return set_error(SMFGPU_ERR_INVALID_ARG, "plan: front rows do not start with the pivots");
        for (int32_t j = D->sn_start[s]; j < D->sn_end[s]; ++j)
            for (int32_t k = D->pcol_ptr[j]; k < D->pcol_ptr[j + 1]; ++k) {
                const int32_t r = pos[D->prow_idx[k]];
                if (r < 0) return set_error(SMFGPU_ERR_INVALID_ARG, "plan: entry outside its front");
                amap[k] = r;
            }
        for (int32_t c = D->child_ptr[s]; c < D->child_ptr[s + 1]; ++c) {
            const int32_t ch = D->child_idx[c];
            const FrontDev &C = P->fronts[ch];
            const int32_t *cr = D->front_row_idx + C.rofs + C.p;
            int32_t prev = -1;
            for (int ii = 0; ii < C.q; ++ii) {
                const int32_t r = pos[cr[ii]];
                if (r < 0 || r <= prev) return set_error(SMFGPU_ERR_INVALID_ARG, "plan: child row outside parent front");
                cmap[cmofs[ch] + ii] = r;
                prev = r;
            }
        }
        for (int t = 0; t < F.f; ++t) pos[fr[t]] = -1;
    }
    // assembly tasks per level, child contributions per parent column
    // (SPD: sizes are static, the whole launch sequence is planned here)
    std::vector<AsmTask> tasks;
    std::vector<ChildCol> cc;
    std::vector<int64_t> asm_first(nlev + 1, 0);
    std::vector<std::vector<ChildCol>> bucket;
    if (P->posdef) {
    for (int32_t lv = 0; lv < nlev; ++lv) {
        asm_first[lv] = static_cast<int64_t>(tasks.size());
        for (int32_t s : P->level_fronts[lv]) {
            const FrontDev &F = P->fronts[s];
            bucket.assign(F.f, {});
            for (int32_t c = D->child_ptr[s]; c < D->child_ptr[s + 1]; ++c) {
                const int32_t ch = D->child_idx[c];
                const FrontDev &C = P->fronts[ch];
                for (int jj = 0; jj < C.q; ++jj) bucket[cmap[cmofs[ch] + jj]].push_back(ChildCol{ch, jj});
            }
            for (int c = 0; c < F.f; ++c) {
                AsmTask t{s, c, static_cast<int32_t>(cc.size()), 0};
                cc.insert(cc.end(), bucket[c].begin(), bucket[c].end());
                t.cc1 = static_cast<int32_t>(cc.size());
                tasks.push_back(t);
            }
        }
    }
    asm_first[nlev] = static_cast<int64_t>(tasks.size());
    if (cc.size() > static_cast<size_t>(INT32_MAX)) return set_error(SMFGPU_ERR_ALLOC, "plan: too many child columns");
    P->asm_first = asm_first;
    }  // posdef
    // value map: slot -> input entries in input order
    const int64_t nslots = P->nnz_perm;
    std::vector<int64_t> slot_ptr(nslots + 1, 0), slot_src;
    for (int64_t e = 0; e < D->n_orig; ++e) {
        const int32_t sl = D->orig_to_slot[e];
        if (sl >= 0) {
            if (sl >= nslots) return set_error(SMFGPU_ERR_INVALID_ARG, "plan: orig_to_slot out of range");
            ++slot_ptr[sl + 1];
        }
    }
    for (int64_t s = 0; s < nslots; ++s) slot_ptr[s + 1] += slot_ptr[s];
    slot_src.resize(slot_ptr[nslots]);
    {
        std::vector<int64_t> fill(slot_ptr.begin(), slot_ptr.end() - 1);
        for (int64_t e = 0; e < D->n_orig; ++e) {
            const int32_t sl = D->orig_to_slot[e];
            if (sl >= 0) slot_src[fill[sl]++] = e;
        }
    }
    // full symmetric CSR for the residual (permuted numbering)
    std::vector<int64_t> fr_ptr(n + 1, 0);
    for (int32_t j = 0; j < n; ++j)
        for (int32_t k = D->pcol_ptr[j]; k < D->pcol_ptr[j + 1]; ++k) {
            const int32_t i = D->prow_idx[k];
            ++fr_ptr[i + 1];
            if (i != j) ++fr_ptr[j + 1];
        }
    for (int32_t i = 0; i < n; ++i) fr_ptr[i + 1] += fr_ptr[i];
    std::vector<int32_t> fr_col(fr_ptr[n]);
    std::vector<int64_t> fr_slot(fr_ptr[n]);
    {
        std::vector<int64_t> fill(fr_ptr.begin(), fr_ptr.end() - 1);
        for (int32_t j = 0; j < n; ++j)
            for (int32_t k = D->pcol_ptr[j]; k < D->pcol_ptr[j + 1]; ++k) {
                const int32_t i = D->prow_idx[k];
                fr_col[fill[i]] = j;
                fr_slot[fill[i]++] = k;
                if (i != j) {
                    fr_col[fill[j]] = i;
                    fr_slot[fill[j]++] = k;
                }
            }
    }
    // uploads
    int rc;
    std::vector<int32_t> rows(D->front_row_idx, D->front_row_idx + P->rows_total);
    std::vector<int32_t> pcp(D->pcol_ptr, D->pcol_ptr + n + 1);
    if ((rc = P->d_fronts.upload(P->fronts)) || (rc = P->d_rows.upload(rows)) || (rc = P->d_cmap.upload(cmap)) ||
        (rc = P->d_cmofs.upload(cmofs)) || (rc = P->d_pcol_ptr.upload(pcp)) || (rc = P->d_amap.upload(amap)) ||
        (rc = P->d_perm.upload(P->perm)) ||
        (rc = P->d_slot_ptr.upload(slot_ptr)) || (rc = P->d_slot_src.upload(slot_src)) ||
        (rc = P->d_asm.upload(tasks)) || (rc = P->d_cc.upload(cc)) || (rc = P->d_fr_ptr.upload(fr_ptr)) ||
        (rc = P->d_fr_col.upload(fr_col)) || (rc = P->d_fr_slot.upload(fr_slot)))
        return rc;
    const bool spd = P->posdef != 0;
    if ((rc = P->d_av.alloc(std::max<int64_t>(nslots, 1))) ||
        (rc = P->d_L.alloc(spd ? std::max<int64_t>(P->factor_entries, 1) : 1)) ||
        (rc = P->d_CB.alloc(spd ? std::max<int64_t>(P->cb_entries, 1) : 1)) || (rc = P->d_red.alloc(8)) ||
        (rc = P->d_flag.alloc(4)) || (rc = P->d_prec.alloc(std::max(ns, 1))) || (rc = P->d_junk.alloc(1024)))
        return rc;
    if (!spd) return ldlt_plan_build(P, D);
    if ((rc = dense_schedule(P))) return rc;
    if ((rc = solve_lists(P))) return rc;
    // algorithmic bytes (assembly, one solve sweep pair per RHS)
    double asm_b = 12.0 * static_cast<double>(P->nnz_perm);
    for (int32_t s = 0; s < ns; ++s) {
        const double f = P->fronts[s].f, q = P->fronts[s].q;
        asm_b += 8.0 * f * (f + 1) / 2.0;
        if (D->sn_parent[s] >= 0) asm_b += 8.0 * q * (q + 1) / 2.0 + 4.0 * q;
    }
    P->asm_bytes = asm_b;
    return SMFGPU_OK;
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" {

SMFGPU_API const char *smfgpu_error_string(void) { return last_error(); }

SMFGPU_API int smfgpu_default_options(smfgpu_factor_options *opt) {
    if (!opt) return set_error(SMFGPU_ERR_INVALID_ARG, "null options");
    std::memset(opt, 0, sizeof(*opt));
    opt->posdef = 1;
    opt->pivot_u = 0.01;
    opt->small_pivot = 1e-20;
    opt->panel = 16;
    opt->precision = SMFGPU_PREC_FP64;
    opt->mixed_min_front = 1024;
    opt->max_negative_pivots = -1;
    opt->value_mem = SMFGPU_MEM_HOST;
    opt->selector_tol = 1e-4;
    return SMFGPU_OK;
}

SMFGPU_API int smfgpu_analyse(int32_t n, const int32_t *col_ptr, const int32_t *row_idx, const int32_t *perm,
                              int32_t nemin, smfgpu_analysis_t *out) {
    if (!out) return set_error(SMFGPU_ERR_INVALID_ARG, "null out");
    *out = nullptr;
    try {
        auto h = std::make_unique<smfgpu_analysis_s>();
        const int rc = analyse(n, col_ptr, row_idx, perm, nemin, h->an);
        if (rc != SMFGPU_OK) return rc;
        *out = h.release();
        return SMFGPU_OK;
    } catch (const std::bad_alloc &) {
        return set_error(SMFGPU_ERR_ALLOC, "analyse: host allocation failed");
    } catch (const std::exception &e) {
        return set_error(SMFGPU_ERR_INVALID_ARG, e.what());
    }
}

SMFGPU_API int smfgpu_analysis_describe(smfgpu_analysis_t an, smfgpu_analysis_desc *desc) {
    if (!an || !desc) return set_error(SMFGPU_ERR_INVALID_ARG, "null handle");
    an->an.describe(*desc);
    return SMFGPU_OK;
}

SMFGPU_API int smfgpu_analysis_stats(smfgpu_analysis_t an, double *out) {
    if (!an || !out) return set_error(SMFGPU_ERR_INVALID_ARG, "null handle");
    const Analysis &a = an->an;
    double rows = static_cast<double>(a.front_row_ptr.empty() ? 0 : a.front_row_ptr.back());
    double fp = 0, fl = 0;
    int maxf = 0;
    std::vector<int> lev(a.nsup, 0);
    int depth = 0;
    for (int s = 0; s < a.nsup; ++s) {
        const int f = a.front_row_ptr[s + 1] - a.front_row_ptr[s];
        const int p = a.sn_end[s] - a.sn_start[s];
        fp += static_cast<double>(f) * p;
        maxf = std::max(maxf, f);
        for (int c = 0; c < p; ++c) {
            const double rem = f - c - 1.0;
            fl += rem + rem * (rem + 1.0);
        }
        if (a.sn_parent[s] >= 0) lev[a.sn_parent[s]] = std::max(lev[a.sn_parent[s]], lev[s] + 1);
        depth = std::max(depth, lev[s] + 1);
    }
    out[0] = a.nsup;
    out[1] = rows;
    out[2] = fp;
    out[3] = maxf;
    out[4] = fl;
    out[5] = depth;
    return SMFGPU_OK;
}

SMFGPU_API void smfgpu_analysis_free(smfgpu_analysis_t *an) {
    if (an && *an) {
        delete *an;
        *an = nullptr;
    }
}

SMFGPU_API int smfgpu_plan_create(const smfgpu_analysis_desc *desc, int32_t posdef, smfgpu_plan_t *out) {
    if (!desc || !out) return set_error(SMFGPU_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return set_error(SMFGPU_ERR_DEVICE, "no HIP device");
    try {
        auto P = std::make_unique<smfgpu_plan_s>();
        P->posdef = posdef;
        HIP_OK(hipGetDevice(&P->device));
        // the no-allocation solve's work block on this device
        if (const int rc = scratch_init(P->device); rc != SMFGPU_OK) return rc;
        HIP_OK(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
        const int rc = plan_build(P.get(), desc);
        if (rc != SMFGPU_OK) {
            (void)hipStreamDestroy(P->stream);
            return rc;
        }
        *out = P.release();
        return SMFGPU_OK;
    } catch (const std::bad_alloc &) {
        return set_error(SMFGPU_ERR_ALLOC, "plan: host allocation failed");
    }
}

SMFGPU_API void smfgpu_plan_free(smfgpu_plan_t *plan) {
    if (!plan || !*plan) return;
    smfgpu_plan_s *P = *plan;
    if (P->stream) {
        (void)hipStreamSynchronize(P->stream);
        (void)hipStreamDestroy(P->stream);
    }
    for (hipEvent_t e : P->ev) (void)hipEventDestroy(e);
    for (hipEvent_t e : P->sev) (void)hipEventDestroy(e);
    release_graph(P);
    // factor handles outlive the plan only as empty shells: every later
    // call on them reports SMFGPU_ERR_STATE
    for (smfgpu_factor_s *f : P->issued) {
        f->plan = nullptr;
        f->valid = false;
    }
    delete P;
    *plan = nullptr;
}

SMFGPU_API int smfgpu_plan_stats(smfgpu_plan_t P, double *out) {
    if (!P || !out) return set_error(SMFGPU_ERR_INVALID_ARG, "null handle");
    double bytes = 0;
    bytes += P->d_L.bytes() + P->d_CB.bytes() + P->d_av.bytes() + P->d_rows.bytes() + P->d_cmap.bytes() +
             P->d_asm.bytes() + P->d_cc.bytes() + P->d_potrf.bytes() + P->d_trsm.bytes() + P->d_syrk.bytes() +
             P->d_amap.bytes() + P->d_slot_src.bytes() + P->d_fr_col.bytes() + P->d_fr_slot.bytes() +
             P->d_fronts.bytes() + P->d_pcol_ptr.bytes() + P->d_perm.bytes() + P->d_level_fronts.bytes() +
             P->d_cmofs.bytes() + P->d_slot_ptr.bytes() + P->d_vals.bytes() + P->d_L32.bytes() +
             P->d_L16h.bytes() + P->d_L16l.bytes() +
             P->d_fwd_items.bytes() + P->d_bwd_items.bytes() + P->d_bwd_seq.bytes() + P->d_fwd_step_of.bytes() +
             P->d_bwd_step_of.bytes() + P->d_fwd_step_cnt.bytes() + P->d_bwd_step_cnt.bytes() +
             P->d_sweep_ctr.bytes() + P->d_sweep_err.bytes() + P->d_gpart.bytes() + P->d_big.bytes() +
             P->d_fr_ptr.bytes();
    // solve and refinement work buffers (allocated on first use)
    bytes += P->d_y.bytes() + P->d_b.bytes() + P->d_r.bytes() + P->d_rowstat.bytes() + P->d_x.bytes() +
             P->d_bb.bytes() + P->d_rvec.bytes() + P->d_stat.bytes() + P->d_part.bytes() + P->d_h.bytes() +
             P->d_av_ref.bytes();
    // LDL^T fronts, factor store and staging
    bytes += P->d_fl.bytes() + P->d_arena.bytes() + P->d_work.bytes() + P->d_dinv.bytes() + P->d_fac.bytes() +
             P->d_lrows.bytes() + P->d_lcmap.bytes() + P->d_flist.bytes() + P->d_ltags.bytes() +
             P->d_dkind.bytes() + P->d_stage.bytes() + P->d_child_ptr.bytes() + P->d_child_idx.bytes() +
             P->d_parent.bytes() + P->d_xpl.bytes();
    out[0] = bytes;
    out[1] = static_cast<double>(P->level_fronts.size());
    out[2] = static_cast<double>(P->launches.size());
    out[3] = P->flops;
    out[4] = static_cast<double>(P->factor_entries);
    out[5] = static_cast<double>(P->cb_entries);
    out[6] = P->asm_bytes;
    double sb = 0;
    for (const FrontDev &F : P->fronts) sb += 2.0 * (8.0 * F.f * static_cast<double>(F.p) + 4.0 * F.f) + 16.0 * F.f;
    out[7] = sb;
    return SMFGPU_OK;
}

SMFGPU_API int64_t smfgpu_launch_profile(smfgpu_plan_t P, double *out, int64_t cap) {
    if (!P) return set_error(SMFGPU_ERR_INVALID_ARG, "null handle");
    const int64_t nl = static_cast<int64_t>(P->launches.size());
    if (!out) return nl;
    const bool have = P->timing && P->ev.size() >= static_cast<size_t>(2 * nl + 2);
    for (int64_t i = 0; i < std::min(nl, cap); ++i) {
        const Launch &l = P->launches[i];
        float ms = -1.0f;
        if (have && (l.kind == K_SYRK || l.kind == K_ASSEMBLE) &&
            hipEventElapsedTime(&ms, P->ev[2 * i], P->ev[2 * i + 1]) != hipSuccess)
            ms = -1.0f;
        out[5 * i + 0] = l.kind;
        out[5 * i + 1] = l.level;
        out[5 * i + 2] = static_cast<double>(l.count);
        out[5 * i + 3] = static_cast<double>(P->syrk_flops_of_launch[i]);
        out[5 * i + 4] = ms;
    }
    return nl;
}

SMFGPU_API int smfgpu_plan_tune(smfgpu_plan_t P, int32_t key, int64_t value) {
    if (!P) return set_error(SMFGPU_ERR_INVALID_ARG, "null handle");
    switch (key) {
    case SMFGPU_TUNE_SOLVE_FUSE_ROWS:
        if (value < 0 || value > (int64_t{1} << 30)) return set_error(SMFGPU_ERR_INVALID_ARG, "tune: bad value");
        P->solve_fuse_rows = static_cast<int32_t>(value);
        return P->posdef ? solve_lists(P) : SMFGPU_OK;
    case SMFGPU_TUNE_PANEL_OUTER:
        if (value < kPanelInner || value % kPanelInner != 0 || value > 4096)
            return set_error(SMFGPU_ERR_INVALID_ARG, "tune: outer panel must be a multiple of 64 in [64, 4096]");
        P->panel_outer = static_cast<int32_t>(value);
        return P->posdef ? dense_schedule(P) : SMFGPU_OK;
    case SMFGPU_TUNE_CB_SEG:
        if (value < 64 || value > (int64_t{1} << 30)) return set_error(SMFGPU_ERR_INVALID_ARG, "tune: bad segment");
        P->cb_seg = static_cast<int32_t>(value);
        return P->posdef ? dense_schedule(P) : SMFGPU_OK;
    case SMFGPU_TUNE_LDLT_BIG_MIN:
        if (value < 96 || value > (int64_t{1} << 30)) return set_error(SMFGPU_ERR_INVALID_ARG, "tune: bad front size");
        P->ldlt_big_min = static_cast<int32_t>(value);
        return SMFGPU_OK;
    case SMFGPU_TUNE_LDLT_GROUP:
        if (value != 0 && value != 1) return set_error(SMFGPU_ERR_INVALID_ARG, "tune: group must be 0 or 1");
        P->ldlt_group = static_cast<int32_t>(value);
        return SMFGPU_OK;
    case SMFGPU_TUNE_LDLT_UPDATE_WG:
        if (value < 8 || value % 8 != 0 || value > (1 << 20))
            return set_error(SMFGPU_ERR_INVALID_ARG, "tune: update workgroups must be a multiple of 8");
        P->ldlt_update_wg = static_cast<int32_t>(value);
        return SMFGPU_OK;
    case SMFGPU_TUNE_LDLT_DEVICE_FRONTS:
        if (value < 0 || value > kDevFrontMax) return set_error(SMFGPU_ERR_INVALID_ARG, "tune: device fronts in [0, 95]");
        P->dev_fmax = static_cast<int32_t>(value);
        return SMFGPU_OK;
    case SMFGPU_TUNE_LDLT_DEVICE_WAVE:
        if (value != 0 && value != 1) return set_error(SMFGPU_ERR_INVALID_ARG, "tune: 0 or 1");
        P->dev_wave = static_cast<int32_t>(value);
        return SMFGPU_OK;
    case SMFGPU_TUNE_ASM_CHUNK:
        if (value < 256 || value > 16384) return set_error(SMFGPU_ERR_INVALID_ARG, "tune: chunk in [256, 16384]");
        P->asm_chunk = static_cast<int32_t>(value);
        release_graph(P);
        return SMFGPU_OK;
    case SMFGPU_TUNE_ASM_WAVE_ROWS:
        if (value < 0 || value > kAsmWaveRowsMax)
            return set_error(SMFGPU_ERR_INVALID_ARG, "tune: wave assembly rows in [0, 4096]");
        P->asm_wave_rows = static_cast<int32_t>(value);
        release_graph(P);
        return SMFGPU_OK;
    case SMFGPU_TUNE_PANEL_OUTER_REDUCED:
        if (value < kPanelInner || value % kPanelInner != 0 || value > 4096)
            return set_error(SMFGPU_ERR_INVALID_ARG, "tune: outer panel must be a multiple of 64 in [64, 4096]");
        P->panel_outer_red = static_cast<int32_t>(value);
        return P->posdef ? dense_schedule(P) : SMFGPU_OK;
    case SMFGPU_TUNE_SOLVE_PERSISTENT:
        if (value != 0 && value != 1) return set_error(SMFGPU_ERR_INVALID_ARG, "tune: 0 or 1");
        P->sweep_persistent = static_cast<int32_t>(value);
        return SMFGPU_OK;
    case SMFGPU_TUNE_LDLT_PLANES:
        if (value != 0 && value != 1) return set_error(SMFGPU_ERR_INVALID_ARG, "tune: 0 or 1");
        P->ldlt_planes = static_cast<int32_t>(value);
        return SMFGPU_OK;
    case SMFGPU_TUNE_LDLT_OUTER_WIDE:
        if (value < 64 || value > kBigOuterMax || value % 64 != 0)
            return set_error(SMFGPU_ERR_INVALID_ARG, "tune: outer block must be a multiple of 64 in [64, 1024]");
        P->ldlt_outer_wide = static_cast<int32_t>(value);
        return SMFGPU_OK;
    case SMFGPU_TUNE_SYRK_AHEAD:
        if (value != 1 && value != 2) return set_error(SMFGPU_ERR_INVALID_ARG, "tune: chunks ahead must be 1 or 2");
        P->syrk_ahead = static_cast<int32_t>(value);
        release_graph(P);
        return SMFGPU_OK;
    default:
        return set_error(SMFGPU_ERR_INVALID_ARG, "tune: unknown key");
    }
}

SMFGPU_API int smfgpu_set_timing(smfgpu_plan_t P, int32_t enabled) {
    if (!P) return set_error(SMFGPU_ERR_INVALID_ARG, "null handle");
    P->timing = enabled != 0;
    return SMFGPU_OK;
}

SMFGPU_API int smfgpu_kernel_times(smfgpu_plan_t P, double *out) {
    if (!P || !out) return set_error(SMFGPU_ERR_INVALID_ARG, "null handle");
    out[0] = P->t_syrk_ms;
    out[1] = P->syrk_launches;
    out[2] = P->syrk_flops;
    out[3] = P->t_asm_ms;
    out[4] = P->asm_bytes;
    out[5] = P->t_solve_ms;
    out[6] = P->solve_bytes;
    for (int k = 0; k < 4; ++k) out[7 + k] = P->t_solve_phase[k];
    out[11] = static_cast<double>(P->dev_levels);
    for (int k = 0; k < 4; ++k) out[12 + k] = P->t_ldlt_host[k] * 1e3;
    return SMFGPU_OK;
}

SMFGPU_API int smfgpu_factor(smfgpu_plan_t P, const double *values, int64_t nvalues,
                             const smfgpu_factor_options *opt_in, smfgpu_factor_t *out, smfgpu_factor_info *info) {
    if (!P || !out) return set_error(SMFGPU_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    smfgpu_factor_options opt;
    if (opt_in) opt = *opt_in;
    else smfgpu_default_options(&opt);
    if ((opt.posdef != 0) != (P->posdef != 0))
        return set_error(SMFGPU_ERR_INVALID_ARG, "factor: matrix type differs from the plan's");
    if (nvalues != P->n_orig) return set_error(SMFGPU_ERR_INVALID_ARG, "factor: wrong number of values");
    const bool autosel = opt.precision == SMFGPU_PREC_AUTO;
    if (!autosel && prec_code(opt.precision) < 0)
        return set_error(SMFGPU_ERR_INVALID_ARG, "factor: precision " + std::to_string(opt.precision) + " is not available");
    if (autosel && !P->sel_loaded)
        return set_error(SMFGPU_ERR_UNSUPPORTED, "factor: precision AUTO needs selector weights (smfgpu_selector_set)");
    if (opt.collect && opt.precision != SMFGPU_PREC_FP64)
        return set_error(SMFGPU_ERR_INVALID_ARG, "factor: collect needs an FP64 factorization");
    if (autosel && !(opt.selector_tol > 0.0))
        return set_error(SMFGPU_ERR_INVALID_ARG, "factor: selector_tol must be > 0");
    if (nvalues > 0 && !values) return set_error(SMFGPU_ERR_INVALID_ARG, "factor: null values");
    hipStream_t st = P->stream;
    // ||A||_inf of the factored values, cached by the refinement, belongs to
    // the previous values
    P->norm_a_valid = false;
    const double *dv = values;
    // values go through d_vals when they are on the host, and for the SPD
    // replayed sequence (its nodes read a fixed buffer)
    const bool via_buf = opt.value_mem == SMFGPU_MEM_HOST || (P->posdef && !P->timing);
    if (via_buf && nvalues > 0) {
        if (P->d_vals.n < static_cast<size_t>(nvalues)) {
            const int rc = P->d_vals.alloc(nvalues);
            if (rc) return rc;
            release_graph(P);
        }
        HIP_OK(hipMemcpyAsync(P->d_vals.p, values, nvalues * sizeof(double),
                              opt.value_mem == SMFGPU_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice,
                              st));
        dv = P->d_vals.p;
    }
    // symmetric scaling: the scale into d_scale, checked, for the value
    // preparation's S A S pass and the solves
    P->scaled = false;
    if (opt.scale) {
        const int32_t n = P->n;
        int rc;
        if (P->d_scale.n < static_cast<size_t>(std::max(n, 1)) || P->d_av_raw.n < static_cast<size_t>(std::max<int64_t>(P->nnz_perm, 1))) {
            if ((rc = P->d_scale.alloc(std::max(n, 1))) || (rc = P->d_av_raw.alloc(std::max<int64_t>(P->nnz_perm, 1))))
                return rc;
            release_graph(P);  // the replayed sequence reads both buffers
        }
        if (opt.scale_mem == SMFGPU_MEM_HOST) {
            for (int32_t i = 0; i < n; ++i)
                if (!(opt.scale[i] > 0.0) || !std::isfinite(opt.scale[i]))
                    return set_error(SMFGPU_ERR_INVALID_ARG, "factor: scale entries must be finite and positive");
            HIP_OK(hipMemcpyAsync(P->d_scale.p, opt.scale, n * sizeof(double), hipMemcpyHostToDevice, st));
        } else {
            HIP_OK(hipMemcpyAsync(P->d_scale.p, opt.scale, n * sizeof(double), hipMemcpyDeviceToDevice, st));
            launch_check_scale(st, P->d_scale.p, n, P->d_red.p + 4);
            unsigned long long bad = 0;
            HIP_OK(hipMemcpyAsync(&bad, P->d_red.p + 4, sizeof(bad), hipMemcpyDeviceToHost, st));
            HIP_OK(hipStreamSynchronize(st));
            if (bad) return set_error(SMFGPU_ERR_INVALID_ARG, "factor: scale entries must be finite and positive");
        }
        opt.scaled = 1;
        P->scaled = true;
    }
    if (!P->posdef) {
        auto F = std::make_unique<smfgpu_factor_s>();
        F->plan = P;
        smfgpu_factor_info fi{};
        if (P->live) P->live->valid = false;
        P->live = nullptr;
        F->auto_prec = autosel;
        F->collected = opt.collect != 0;
        F->mixed_min_front = opt.mixed_min_front;
        if (autosel || opt.collect) {
            if ((P->d_feat.p == nullptr && P->d_feat.alloc(static_cast<size_t>(std::max(P->nsup, 1)) * kSelNF)))
                return SMFGPU_ERR_ALLOC;
            if (opt.collect) {
                if (P->d_lab.p == nullptr && P->d_lab.alloc(static_cast<size_t>(std::max(P->nsup, 1)) * 4))
                    return SMFGPU_ERR_ALLOC;
                HIP_OK(hipMemsetAsync(P->d_lab.p, 0, static_cast<size_t>(P->nsup) * 4 * sizeof(unsigned int), st));
            }
        }
        const double t0 = now_s();
        const int rc = ldlt_factor(P, dv, nvalues, opt, fi);
        fi.factor_seconds = now_s() - t0;
        if (info) *info = fi;
        if (rc != SMFGPU_OK) return rc;
        F->prec = P->ldlt_prec;
        F->info = fi;
        // solve_core refuses a singular factor (solve.cpp:888-903)
        F->valid = fi.status == SMFGPU_STATUS_SUCCESS;
        P->live = F.get();
        P->issued.insert(F.get());
        P->max_abs = fi.max_abs;
        *out = F.release();
        return SMFGPU_OK;
    }
    // per-front Schur precision: fronts of at least mixed_min_front rows
    // run the requested reduced precision, or, with AUTO, the precision the
    // learned selector picks on the device once the front is assembled
    auto F = std::make_unique<smfgpu_factor_s>();
    F->plan = P;
    F->prec.assign(P->nsup, 0);
    F->auto_prec = autosel;
    F->collected = opt.collect != 0;
    F->mixed_min_front = opt.mixed_min_front;
    int reduced = 0;
    const int8_t code = autosel ? 0 : static_cast<int8_t>(prec_code(opt.precision));
    int candidates = 0;
    for (int32_t s = 0; s < P->nsup; ++s)
        if (P->fronts[s].f >= opt.mixed_min_front) {
            ++candidates;
            if (code > 0) {
                F->prec[s] = code;
                ++reduced;
            }
        }
    const bool sel_on = (autosel || opt.collect) && candidates > 0;
    const int gi = autosel && candidates > 0 ? 2 : (reduced > 0 ? 1 : 0);
    if (sel_on) {
        if ((P->d_colstat.p == nullptr && P->d_colstat.alloc(std::max(P->n, 1))) ||
            (P->d_feat.p == nullptr && P->d_feat.alloc(static_cast<size_t>(std::max(P->nsup, 1)) * kSelNF)))
            return SMFGPU_ERR_ALLOC;
        // the graph of the selector's launches bakes its parameters in
        const double key[3] = {opt.selector_tol, static_cast<double>(opt.selector_mask), static_cast<double>(opt.scaled)};
        if (autosel && (key[0] != P->sel_key[0] || key[1] != P->sel_key[1] || key[2] != P->sel_key[2])) {
            release_graph(P);
            std::copy(key, key + 3, P->sel_key);
        }
    }
    if ((reduced > 0 || sel_on) && P->sched_red_min != opt.mixed_min_front) {
        // the Schur launches put the items of fronts that may run reduced last
        P->sched_red_min = std::max(opt.mixed_min_front, 1);
        const int rc = dense_schedule(P);
        if (rc) return rc;
    }
    const bool need16 = (autosel && candidates > 0) || (reduced > 0 && (code == 2 || code == 3));
    const bool need16lo = (autosel && candidates > 0) || (reduced > 0 && code == 3);
    if ((need16 && P->d_L16h.p == nullptr) || (need16lo && P->d_L16l.p == nullptr)) {
        // packed bf16 panel planes (the BF16 tiles' operands); zeros in the
        // padding columns past each front's last pivot stay zero
        for (DBuf<uint16_t> *pb : {&P->d_L16h, &P->d_L16l}) {
            const bool want = pb == &P->d_L16h ? need16 : need16lo;
            if (!want || pb->p != nullptr) continue;
            if (const int rc = pb->alloc(std::max<int64_t>(P->l16_entries, 1))) return rc;
            HIP_OK(hipMemsetAsync(pb->p, 0, pb->bytes(), st));
        }
        release_graph(P);
    }
    if ((reduced > 0 || (autosel && candidates > 0)) && P->d_L32.p == nullptr) {
        // FP32 panel copies (the reduced fronts' Schur operands), once per plan
        int rc = P->d_L32.alloc(std::max<int64_t>(P->factor_entries, 1));
        if (rc) return rc;
        release_graph(P);
    }
    HIP_OK(hipMemcpyAsync(P->d_prec.p, F->prec.data(), P->nsup, hipMemcpyHostToDevice, st));
    const int32_t big = INT32_MAX;
    HIP_OK(hipMemcpyAsync(P->d_flag.p, &big, sizeof(int32_t), hipMemcpyHostToDevice, st));
    DevPtrs d = dev_ptrs(P);
    SelPrm sprm{};
    if (sel_on) {
        d.colstat = P->d_colstat.p;
        sprm = sel_prm(P, opt);
    }
    // reduced-precision Schur kernels to launch: the static code, or every
    // code the selector may pick
    const int red_modes = autosel ? sprm.mask : (code > 0 ? (1 << code) : 0);
    // the launch sequence of this factorization's precision class
    const std::vector<Launch> &LL = gi ? P->launches_red : P->launches;
    const std::vector<int64_t> &LF = gi ? P->syrk_flops_of_launch_red : P->syrk_flops_of_launch;
    // timing events
    const bool timing = P->timing;
    const size_t nl = LL.size();
    if (timing && P->ev.size() < 2 * nl + 2) {
        for (size_t i = P->ev.size(); i < 2 * nl + 2; ++i) {
            hipEvent_t e;
            HIP_OK(hipEventCreate(&e));
            P->ev.push_back(e);
        }
    }
    const double t0 = now_s();
    auto enqueue = [&](bool rec_all) -> hipError_t {
        hipError_t e = hipSuccess;
        auto mark = [&](size_t i) {
            if (e == hipSuccess) e = hipEventRecord(P->ev[i], st);
        };
        if (rec_all) mark(2 * nl);
        launch_map_values(st, dv, nvalues, P->d_slot_ptr.p, P->d_slot_src.p, P->nnz_perm, P->d_av.p, P->d_red.p);
        if (P->scaled)
            launch_scale_values(st, P->n, P->d_fr_ptr.p, P->d_fr_col.p, P->d_fr_slot.p, P->d_perm.p, P->d_scale.p,
                                P->d_av.p, P->d_av_raw.p, P->d_red.p);
        for (size_t li = 0; li < nl; ++li) {
            const Launch &l = LL[li];
            const bool rec = rec_all && (l.kind == K_SYRK || l.kind == K_ASSEMBLE);
            if (rec) mark(2 * li);
            switch (l.kind) {
            case K_ASSEMBLE: {
                int32_t fmax = 1;
                for (int32_t s : P->level_fronts[l.level]) fmax = std::max(fmax, P->fronts[s].f);
                launch_assemble(st, d, P->d_asm.p + l.first, l.count, P->d_cc.p, fmax, P->asm_chunk, P->asm_wave_rows);
                break;
            }
            case K_POTRF: launch_potrf(st, d, P->d_potrf.p + l.first, l.count); break;
            case K_TRSM: launch_trsm(st, d, P->d_trsm.p + l.first, l.count); break;
            case K_SYRK:
                launch_syrk(st, d, P->d_syrk.p + l.first, l.count, gi ? l.red : 0, P->syrk_ahead, red_modes);
                break;
            case K_SELECT:
                if (sel_on) launch_select(st, P->d_selfr.p + l.first, static_cast<int32_t>(l.count), P->d_colstat.p, sprm,
                                          P->d_feat.p, P->d_prec.p);
                break;
            default: break;
            }
            if (rec) mark(2 * li + 1);
        }
        if (rec_all) mark(2 * nl + 1);
        const hipError_t e2 = hipGetLastError();
        return e != hipSuccess ? e : e2;
    };
    if (timing || nvalues == 0 || opt.collect) {
        HIP_OK(enqueue(timing));
    } else {
        // the captured sequence: its reduced-precision kernels, and the
        // scaling pass when the values are scaled
        const int32_t gkey = red_modes | (P->scaled ? (1 << 20) : 0);
        if (P->fexec[gi] && P->graph_red_modes[gi] != gkey) {
            // captured with another set of kernels
            (void)hipGraphExecDestroy(P->fexec[gi]);
            (void)hipGraphDestroy(P->fgraph[gi]);
            P->fexec[gi] = nullptr;
            P->fgraph[gi] = nullptr;
        }
        if (!P->fexec[gi]) {
            P->graph_red_modes[gi] = gkey;
            HIP_OK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            const hipError_t e1 = enqueue(false);
            hipGraph_t g = nullptr;
            const hipError_t e2 = hipStreamEndCapture(st, &g);
            if (e1 != hipSuccess || e2 != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return set_error(SMFGPU_ERR_DEVICE, "factor: capture of the launch sequence failed");
            }
            P->fgraph[gi] = g;
            if (hipGraphInstantiate(&P->fexec[gi], g, nullptr, nullptr, 0) != hipSuccess) {
                release_graph(P);
                return set_error(SMFGPU_ERR_DEVICE, "factor: graph instantiation failed");
            }
        }
        HIP_OK(hipGraphLaunch(P->fexec[gi], st));
    }
    int32_t flag = 0;
    unsigned long long maxbits = 0;
    if (autosel && candidates > 0)
        HIP_OK(hipMemcpyAsync(F->prec.data(), P->d_prec.p, P->nsup, hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(&flag, P->d_flag.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(&maxbits, P->d_red.p, sizeof(maxbits), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    const double t1 = now_s();
    double max_abs;
    std::memcpy(&max_abs, &maxbits, sizeof(double));
    smfgpu_factor_info fi{};
    fi.max_abs = max_abs;
    fi.max_front = 0;
    for (const FrontDev &Fd : P->fronts) fi.max_front = std::max(fi.max_front, Fd.f);
    if (autosel) {
        reduced = 0;
        for (int8_t c : F->prec) reduced += c != 0;
    }
    fi.fronts_reduced = reduced;
    fi.factor_entries = P->factor_entries;
    fi.factor_seconds = t1 - t0;
    if (timing) {
        float ms = 0;
        HIP_OK(hipEventElapsedTime(&ms, P->ev[2 * nl], P->ev[2 * nl + 1]));
        fi.factor_seconds = ms * 1e-3;
        P->t_syrk_ms = P->t_asm_ms = 0;
        P->syrk_launches = 0;
        P->syrk_flops = 0;
        for (size_t li = 0; li < nl; ++li) {
            const Launch &l = LL[li];
            if (l.kind != K_SYRK && l.kind != K_ASSEMBLE) continue;
            HIP_OK(hipEventElapsedTime(&ms, P->ev[2 * li], P->ev[2 * li + 1]));
            if (l.kind == K_SYRK) {
                P->t_syrk_ms += ms;
                P->syrk_launches += 1;
                P->syrk_flops += static_cast<double>(LF[li]);
            } else {
                P->t_asm_ms += ms;
            }
        }
    }
    if (!(maxbits < 0x7ff0000000000000ull)) {
        fi.status = SMFGPU_STATUS_SINGULAR;
        if (info) *info = fi;
        return set_error(SMFGPU_ERR_VALUE, "factor: non-finite input value");
    }
    int32_t npos = P->n;
    if (flag != INT32_MAX) {
        fi.status = SMFGPU_STATUS_NOT_POSDEF;
        // pivots accepted before the failing front (postorder): those fronts' widths
        npos = 0;
        for (int32_t s = 0; s < flag - 1; ++s) npos += P->fronts[s].p;
    } else {
        fi.status = SMFGPU_STATUS_SUCCESS;
    }
    fi.num_positive = npos;
    fi.actual_flops = P->flops;
    double fw = 0;
    for (const FrontDev &Fd : P->fronts) fw += static_cast<double>(Fd.p) * Fd.f * static_cast<double>(Fd.f);
    fi.front_work = fw;
    F->info = fi;
    F->valid = fi.status == SMFGPU_STATUS_SUCCESS;
    if (P->live) P->live->valid = false;
    P->live = F.get();
    P->issued.insert(F.get());
    P->max_abs = max_abs;
    if (info) *info = fi;
    *out = F.release();
    return SMFGPU_OK;
}

SMFGPU_API int smfgpu_equilibration_scale(smfgpu_plan_t P, const double *values, int64_t nvalues, int32_t value_mem,
                                          int32_t iters, double *scale, int32_t scale_mem) {
    if (!P || !scale) return set_error(SMFGPU_ERR_INVALID_ARG, "null argument");
    if (nvalues != P->n_orig) return set_error(SMFGPU_ERR_INVALID_ARG, "equilibration: wrong number of values");
    if (nvalues > 0 && !values) return set_error(SMFGPU_ERR_INVALID_ARG, "equilibration: null values");
    if (iters < 0) return set_error(SMFGPU_ERR_INVALID_ARG, "equilibration: iters must be >= 0");
    hipStream_t st = P->stream;
    const int32_t n = P->n;
    int rc;
    // the values mapped into the refinement's scratch slots (rewritten by
    // every refinement against other values, so free between calls)
    if (P->d_av_ref.n < static_cast<size_t>(std::max<int64_t>(P->nnz_perm, 1)) &&
        (rc = P->d_av_ref.alloc(std::max<int64_t>(P->nnz_perm, 1))))
        return rc;
    if (P->d_eqs.n < 2 * static_cast<size_t>(std::max(n, 1)) && (rc = P->d_eqs.alloc(2 * static_cast<size_t>(std::max(n, 1)))))
        return rc;
    const double *vd = values;
    if (value_mem == SMFGPU_MEM_HOST && nvalues > 0) {
        if (P->d_vals.n < static_cast<size_t>(nvalues)) {
            if ((rc = P->d_vals.alloc(nvalues))) return rc;
            release_graph(P);
        }
        HIP_OK(hipMemcpyAsync(P->d_vals.p, values, nvalues * sizeof(double), hipMemcpyHostToDevice, st));
        vd = P->d_vals.p;
    }
    launch_map_values(st, vd, nvalues, P->d_slot_ptr.p, P->d_slot_src.p, P->nnz_perm, P->d_av_ref.p, P->d_red.p + 5);
    double *res = nullptr;
    launch_equilibrate(st, n, P->d_fr_ptr.p, P->d_fr_col.p, P->d_fr_slot.p, P->d_perm.p, P->d_av_ref.p, iters,
                       P->d_eqs.p, P->d_eqs.p + std::max(n, 1), &res);
    HIP_OK(hipGetLastError());
    unsigned long long maxbits = 0;
    HIP_OK(hipMemcpyAsync(&maxbits, P->d_red.p + 5, sizeof(maxbits), hipMemcpyDeviceToHost, st));
    if (n > 0)
        HIP_OK(hipMemcpyAsync(scale, res, n * sizeof(double),
                              scale_mem == SMFGPU_MEM_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, st));
    HIP_OK(hipStreamSynchronize(st));
    if (!(maxbits < 0x7ff0000000000000ull)) return set_error(SMFGPU_ERR_VALUE, "equilibration: non-finite input value");
    return SMFGPU_OK;
}

SMFGPU_API void smfgpu_factor_free(smfgpu_factor_t *f) {
    if (!f || !*f) return;
    if ((*f)->plan) {
        if ((*f)->plan->live == *f) (*f)->plan->live = nullptr;
        (*f)->plan->issued.erase(*f);
    }
    delete *f;
    *f = nullptr;
}

SMFGPU_API int smfgpu_factor_front_stats(smfgpu_factor_t f, int32_t *out) {
    if (!f || !out) return set_error(SMFGPU_ERR_INVALID_ARG, "null argument");
    if (!f->plan) return set_error(SMFGPU_ERR_STATE, "front stats: the factor's plan was freed");
    smfgpu_plan_s *P = f->plan;
    if (P->live != f) return set_error(SMFGPU_ERR_STATE, "front stats: factor was superseded by a newer factorization");
    for (int32_t s = 0; s < P->nsup; ++s) {
        int32_t *o = out + 8 * static_cast<int64_t>(s);
        if (P->posdef) {
            const FrontDev &F = P->fronts[s];
            o[0] = F.f; o[1] = F.p; o[2] = F.p; o[3] = 0; o[4] = 0; o[5] = 0; o[6] = 0; o[7] = F.p;
        } else {
            const FrontLd &F = P->fl[s];
            o[0] = F.f; o[1] = F.p; o[2] = F.npos; o[3] = F.nneg; o[4] = F.nzero; o[5] = F.n2x2;
            o[6] = F.delayed; o[7] = F.ne;
        }
    }
    return SMFGPU_OK;
}

SMFGPU_API int smfgpu_factor_stall_stats(smfgpu_factor_t f, int32_t *out) {
    if (!f || !out) return set_error(SMFGPU_ERR_INVALID_ARG, "null argument");
    if (!f->plan) return set_error(SMFGPU_ERR_STATE, "stall stats: the factor's plan was freed");
    smfgpu_plan_s *P = f->plan;
    if (P->live != f) return set_error(SMFGPU_ERR_STATE, "stall stats: factor was superseded by a newer factorization");
    for (int32_t s = 0; s < P->nsup; ++s) {
        int32_t *o = out + 4 * static_cast<int64_t>(s);
        if (P->posdef) {
            o[0] = o[1] = o[2] = o[3] = 0;
        } else {
            const FrontLd &F = P->fl[s];
            o[0] = F.stalls; o[1] = F.pulls; o[2] = F.rotated; o[3] = F.searches;
        }
    }
    return SMFGPU_OK;
}

SMFGPU_API int smfgpu_export_factor(smfgpu_factor_t f, int64_t *sizes, int64_t *steps, double *factor_values,
                                    int32_t *row_indices, int8_t *pivot_types) {
    if (!f) return set_error(SMFGPU_ERR_INVALID_ARG, "null factor");
    if (!f->plan) return set_error(SMFGPU_ERR_STATE, "export: the factor's plan was freed");
    smfgpu_plan_s *P = f->plan;
    if (P->live != f) return set_error(SMFGPU_ERR_STATE, "export: factor was superseded by a newer factorization");
    if (!P->posdef) return ldlt_export(P, sizes, steps, factor_values, row_indices, pivot_types);
    if (sizes) {
        sizes[0] = P->nsup;
        sizes[1] = P->factor_entries;
        sizes[2] = P->rows_total;
        sizes[3] = P->n;
    }
    if (steps)
        for (int32_t s = 0; s < P->nsup; ++s) {
            const FrontDev &F = P->fronts[s];
            steps[4 * s + 0] = F.rofs;
            steps[4 * s + 1] = F.lofs;
            steps[4 * s + 2] = F.p;
            steps[4 * s + 3] = F.f;
        }
    if (factor_values) {
        HIP_OK(hipMemcpy(factor_values, P->d_L.p, P->factor_entries * sizeof(double), hipMemcpyDeviceToHost));
        // the device keeps inv(L_kk)^T in the strict upper part of each
        // pivot block; the stored factor has zeros there
        for (int32_t s = 0; s < P->nsup; ++s) {
            const FrontDev &F = P->fronts[s];
            double *Lb = factor_values + F.lofs;
            for (int c = 0; c < F.p; ++c) std::fill_n(Lb + static_cast<int64_t>(c) * F.f, c, 0.0);
        }
    }
    if (row_indices)
        HIP_OK(hipMemcpy(row_indices, P->d_rows.p, P->rows_total * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (pivot_types) std::fill_n(pivot_types, P->n, static_cast<int8_t>(1));
    return SMFGPU_OK;
}

}  // extern "C"

namespace smfgpu {
// job as ldlt_solve; for a Cholesky factor D = I (the diagonal job copies)
int sweep_check(smfgpu_plan_s *P) {
    if (!P->d_sweep_err.p || !P->h_sweep_err.p) return SMFGPU_OK;
    HIP_OK(hipMemcpyAsync(P->h_sweep_err.p, P->d_sweep_err.p, sizeof(int32_t), hipMemcpyDeviceToHost, P->stream));
    HIP_OK(hipStreamSynchronize(P->stream));
    if (P->h_sweep_err.p[0] == 0) return SMFGPU_OK;
    HIP_OK(hipMemsetAsync(P->d_sweep_err.p, 0, sizeof(int32_t), P->stream));
    HIP_OK(hipStreamSynchronize(P->stream));
    return set_error(SMFGPU_ERR_DEVICE, "solve: a sweep's dependency wait ran out");
}

int solve_device(smfgpu_plan_s *P, const double *b_dev, int64_t ldb, double *out_dev, int64_t ldo, int32_t nrhs,
                 int32_t job, double *ybuf, int64_t ycap) {
    if (!P->posdef) return ldlt_solve(P, b_dev, ldb, out_dev, ldo, nrhs, job, ybuf, ycap);
    hipStream_t st = P->stream;
    const int32_t n = P->n;
    DevPtrs d = dev_ptrs(P);
    const int32_t nl = static_cast<int32_t>(P->level_fronts.size());
    // right-hand sides per pass through the sweeps: 8 (a step's item keeps
    // them in LDS, one wave each in the triangles), or what a caller-given
    // work block holds
    constexpr int32_t kGroup = 8;
    const int32_t group = ybuf ? static_cast<int32_t>(std::min<int64_t>(kGroup, ycap / std::max(n, 1))) : kGroup;
    if (group < 1) return set_error(SMFGPU_ERR_INVALID_ARG, "solve: work block smaller than one column");
    for (int32_t r0 = 0; r0 < nrhs; r0 += group) {
        const int32_t nr = std::min(group, nrhs - r0);
        if (!ybuf && P->d_y.n < static_cast<size_t>(n) * nr) {
            const int rc = P->d_y.alloc(static_cast<size_t>(n) * kGroup);
            if (rc) return rc;
        }
        double *y = ybuf ? ybuf : P->d_y.p;
        const bool fwd = job == 0 || job == 1;
        const double *sc = P->scaled ? P->d_scale.p : nullptr;
        if (fwd) launch_gather_rhs(st, P->d_perm.p, b_dev + r0 * ldb, ldb, y, n, nr, sc);
        else
            HIP_OK(hipMemcpy2DAsync(y, n * sizeof(double), b_dev + r0 * ldb, ldb * sizeof(double),
                                    n * sizeof(double), nr, hipMemcpyDeviceToDevice, st));
        const bool tm = P->timing && job == 0 && !ybuf;
        if (tm && P->sev.size() < 3) {
            for (size_t i = P->sev.size(); i < 3; ++i) {
                hipEvent_t e;
                HIP_OK(hipEventCreate(&e));
                P->sev.push_back(e);
            }
        }
        if (tm) HIP_OK(hipEventRecord(P->sev[0], st));
        const bool pers = P->sweep_persistent != 0;
        if (fwd && pers) {
            launch_sweep(st, d, true, P->d_fwd_items.p, P->d_fwd_step_of.p, P->d_fwd_step_cnt.p,
                         static_cast<int32_t>(P->d_fwd_items.n), static_cast<int32_t>(P->fwd_steps.size()),
                         P->d_sweep_ctr.p, P->d_sweep_err.p, y, n, nr);
        } else if (fwd) {
            for (int32_t lv = 0; lv < nl; ++lv)
                for (int32_t i = P->fwd_lv[lv]; i < P->fwd_lv[lv + 1]; ++i)
                    launch_fwd_step(st, d, P->d_fwd_items.p + P->fwd_steps[i].first, P->fwd_steps[i].second, y,
                                    n, nr);
        }
        if (tm) HIP_OK(hipEventRecord(P->sev[1], st));
        if (job == 1 || job == 2) {
            HIP_OK(hipMemcpy2DAsync(out_dev + r0 * ldo, ldo * sizeof(double), y, n * sizeof(double),
                                    n * sizeof(double), nr, hipMemcpyDeviceToDevice, st));
            continue;
        }
        if (pers) {
            launch_sweep(st, d, false, P->d_bwd_seq.p, P->d_bwd_step_of.p, P->d_bwd_step_cnt.p,
                         static_cast<int32_t>(P->d_bwd_seq.n), static_cast<int32_t>(P->bwd_steps.size()),
                         P->d_sweep_ctr.p, P->d_sweep_err.p, y, n, nr);
        } else {
            for (int32_t lv = nl - 1; lv >= 0; --lv)
                for (int32_t i = P->bwd_lv[lv]; i < P->bwd_lv[lv + 1]; ++i)
                    launch_bwd_step(st, d, P->d_bwd_items.p + P->bwd_steps[i].first, P->bwd_steps[i].second, y,
                                    n, nr);
        }
        if (tm) {
            HIP_OK(hipEventRecord(P->sev[2], st));
            HIP_OK(hipStreamSynchronize(st));
            float ms;
            HIP_OK(hipEventElapsedTime(&ms, P->sev[0], P->sev[1]));
            P->t_solve_phase[0] = ms;
            HIP_OK(hipEventElapsedTime(&ms, P->sev[1], P->sev[2]));
            P->t_solve_phase[1] = ms;
            P->t_solve_phase[2] = static_cast<double>(P->fwd_steps.size());
            P->t_solve_phase[3] = static_cast<double>(P->bwd_steps.size());
        }
        launch_scatter_rhs(st, P->d_perm.p, y, out_dev + r0 * ldo, ldo, n, nr, sc);
    }
    HIP_OK(hipGetLastError());
    return SMFGPU_OK;
}

}  // namespace smfgpu

extern "C" {

SMFGPU_API int smfgpu_solve(smfgpu_factor_t f, double *b, int32_t nrhs, int32_t ldb, int32_t job, int32_t mem) {
    if (!f || !b) return set_error(SMFGPU_ERR_INVALID_ARG, "null argument");
    if (!f->plan) return set_error(SMFGPU_ERR_STATE, "solve: the factor's plan was freed");
    smfgpu_plan_s *P = f->plan;
    if (P->live != f || !f->valid) return set_error(SMFGPU_ERR_STATE, "solve: factor not usable");
    if (job < 0 || job > 4) return set_error(SMFGPU_ERR_INVALID_ARG, "solve: job must be 0..4");
    if (nrhs <= 0 || ldb < P->n) return set_error(SMFGPU_ERR_INVALID_ARG, "solve: bad nrhs/ldb");
    hipStream_t st = P->stream;
    const int32_t n = P->n;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (P->timing) {
        HIP_OK(hipEventCreate(&e0));
        HIP_OK(hipEventCreate(&e1));
    }
    int rc;
    if (mem == SMFGPU_MEM_DEVICE) {
        if (P->timing) HIP_OK(hipEventRecord(e0, st));
        rc = solve_device(P, b, ldb, b, ldb, nrhs, job);
        if (P->timing) HIP_OK(hipEventRecord(e1, st));
    } else {
        if (P->d_b.n < static_cast<size_t>(n) * nrhs && (rc = P->d_b.alloc(static_cast<size_t>(n) * nrhs))) return rc;
        HIP_OK(hipMemcpy2DAsync(P->d_b.p, n * sizeof(double), b, ldb * sizeof(double), n * sizeof(double), nrhs,
                                hipMemcpyHostToDevice, st));
        if (P->timing) HIP_OK(hipEventRecord(e0, st));
        rc = solve_device(P, P->d_b.p, n, P->d_b.p, n, nrhs, job);
        if (P->timing) HIP_OK(hipEventRecord(e1, st));
        HIP_OK(hipMemcpy2DAsync(b, ldb * sizeof(double), P->d_b.p, n * sizeof(double), n * sizeof(double), nrhs,
                                hipMemcpyDeviceToHost, st));
    }
    HIP_OK(hipStreamSynchronize(st));
    if (rc == SMFGPU_OK) rc = sweep_check(P);
    if (P->timing) {
        float ms = 0;
        HIP_OK(hipEventElapsedTime(&ms, e0, e1));
        P->t_solve_ms = ms;
        double sb = 0;
        for (const FrontDev &F : P->fronts) sb += 2.0 * (8.0 * F.f * static_cast<double>(F.p) + 4.0 * F.f) + 16.0 * F.f * nrhs;
        P->solve_bytes = sb;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    return rc;
}

}  // extern "C"
