// Host driver of the indefinite (LDL^T) numeric phase: level by level, the
// host lays out the level's fronts from its children's results (delayed
// columns first, as process_front, solver/src/multifrontal_serial.cpp:510-545
// of the reference), uploads the row lists, child maps and assembly tasks,
// launches assembly, the front kernels and the epilogue, and reads back the
// pivot counts and final row orders the next level needs.
#include "smfgpu_plan.hpp"

#include <algorithm>
#include <cmath>
#include <chrono>
#include <cstring>

namespace smfgpu {

void launch_front_gather(hipStream_t st, const FrontLd *fronts, FrontLd *out, const int32_t *ids, int32_t np);

constexpr int kBlockedMinFront = 96;  // ldlt_kernel.hpp:15 (blocked_min_front)
constexpr int64_t kStoreWide = int64_t{1} << 20;  // stored entries from which a front gets the store kernel

namespace {

double wall_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <class T>
int ensure(DBuf<T> &b, size_t need, bool keep, const char *what = "buffer") {
    if (b.n >= need) return SMFGPU_OK;
    size_t cap = std::max(need, b.n + b.n / 2);
    if (!keep) {
        const int rc = b.alloc(cap);
        return rc ? set_error(rc, std::string("factor: ") + what + ": " + last_error()) : SMFGPU_OK;
    }
    DBuf<T> nb;
    int rc = nb.alloc(cap);
    if (rc) return set_error(rc, std::string("factor: ") + what + ": " + last_error());
    if (b.n && hipMemcpy(nb.p, b.p, b.n * sizeof(T), hipMemcpyDeviceToDevice) != hipSuccess)
        return set_error(SMFGPU_ERR_DEVICE, "arena growth copy failed");
    std::swap(b.p, nb.p);
    std::swap(b.n, nb.n);
    return SMFGPU_OK;
}

LdltDev ldlt_dev(smfgpu_plan_s *P) {
    LdltDev d{};
    d.fronts = P->d_fl.p;
    d.arena = P->d_arena.p;
    d.fac = P->d_fac.p;
    d.rows = P->d_lrows.p;
    d.tags = P->d_ltags.p;
    d.work = P->d_work.p;
    d.dinv = P->d_dinv.p;
    d.dkind = P->d_dkind.p;
    d.cmap = P->d_lcmap.p;
    d.pcol_ptr = P->d_pcol_ptr.p;
    d.amap = P->d_amap.p;
    d.av = P->d_av.p;
    d.junk = P->d_junk.p;
    return d;
}

// staging layout helper: appends a vector's bytes (8-aligned), returns offset
template <class T>
size_t stage(std::vector<unsigned char> &buf, const std::vector<T> &v) {
    size_t off = (buf.size() + 7) & ~size_t{7};
    buf.resize(off + v.size() * sizeof(T));
    if (!v.empty()) std::memcpy(buf.data() + off, v.data(), v.size() * sizeof(T));
    return off;
}

// Peak of the square pool over a factorization without delays: each level
// takes its fronts' squares, then gives back its children's.
int64_t square_peak(const smfgpu_plan_s *P) {
    RangePool pool;
    std::vector<int64_t> off(P->fronts.size(), 0);
    for (const auto &fr : P->level_fronts) {
        for (int32_t s : fr) {
            const int64_t f = P->fronts[s].f;
            off[s] = pool.take(f * f);
        }
        for (int32_t s : fr)
            for (int32_t c = P->child_ptr[s]; c < P->child_ptr[s + 1]; ++c) {
                const int32_t ch = P->child_idx[c];
                const int64_t fc = P->fronts[ch].f;
                pool.give(off[ch], fc * fc);
            }
    }
    return pool.end;
}

}  // namespace

int ldlt_plan_build(smfgpu_plan_s *P, const smfgpu_analysis_desc *D) {
    const int32_t ns = D->nsup;
    P->sn_parent.assign(D->sn_parent, D->sn_parent + ns);
    P->child_ptr.assign(D->child_ptr, D->child_ptr + ns + 1);
    P->child_idx.assign(D->child_idx, D->child_idx + P->child_ptr[ns]);
    P->front_row_ptr.assign(D->front_row_ptr, D->front_row_ptr + ns + 1);
    P->front_row_idx.assign(D->front_row_idx, D->front_row_idx + P->front_row_ptr[ns]);
    int rc;
    if ((rc = P->fl.assign(ns, FrontLd{}))) return rc;
    // level front lists (all fronts) for the solves
    std::vector<int32_t> lf;
    const int32_t nlev = static_cast<int32_t>(P->level_fronts.size());
    P->lv_flist_ofs.assign(nlev + 1, 0);
    for (int32_t lv = 0; lv < nlev; ++lv) {
        P->lv_flist_ofs[lv] = static_cast<int64_t>(lf.size());
        lf.insert(lf.end(), P->level_fronts[lv].begin(), P->level_fronts[lv].end());
    }
    P->lv_flist_ofs[nlev] = static_cast<int64_t>(lf.size());
    // room a device level may need: squares, factor store, rows and D^-1
    // entries of its fronts with up to 8 delayed columns each (a level that
    // needs more is retried with more room)
    P->lv_fsym_max.assign(nlev, 0);
    P->lv_room.assign(4 * static_cast<size_t>(nlev + 1), 0);
    for (int32_t lv = 0; lv < nlev; ++lv) {
        int64_t *r = &P->lv_room[4 * static_cast<size_t>(lv + 1)];
        for (int k = 0; k < 4; ++k) r[k] = r[k - 4];
        for (int32_t s : P->level_fronts[lv]) {
            P->lv_fsym_max[lv] = std::max(P->lv_fsym_max[lv], P->fronts[s].f);
            const int64_t fe = std::min<int64_t>(kDevFrontMax, P->fronts[s].f + 8);
            const int64_t pe = std::min<int64_t>(fe, P->fronts[s].p + 8);
            r[0] += RangePool::round(fe * fe);
            r[1] += fe * pe;
            r[2] += fe;
            r[3] += pe;
        }
    }
    if ((rc = P->d_flist.upload(lf)) || (rc = P->d_fl.alloc(std::max(ns, 1))) ||
        (rc = P->d_child_ptr.upload(P->child_ptr)) || (rc = P->d_child_idx.upload(P->child_idx)) ||
        (rc = P->d_parent.upload(P->sn_parent)) || (rc = P->d_dctr.alloc(kDcN)))
        return rc;
    // initial capacities: the pool's peak when the levels are laid out in
    // order without delays, and the symbolic factor, each with an eighth
    // more for delays
    int64_t fac = 0, rw = 0;
    for (int32_t s = 0; s < ns; ++s) {
        fac += static_cast<int64_t>(P->fronts[s].f) * P->fronts[s].p;
        rw += P->fronts[s].f;
    }
    const int64_t sq_peak = square_peak(P);
    if ((rc = P->d_arena.alloc(static_cast<size_t>(sq_peak + sq_peak / 8 + 1024))) ||
        (rc = P->d_fac.alloc(static_cast<size_t>(fac + fac / 8 + 1024))) ||
        (rc = P->d_lrows.alloc(static_cast<size_t>(rw + rw / 4 + 1024))) ||
        (rc = P->d_ltags.alloc(static_cast<size_t>(rw + rw / 4 + 1024))) ||
        (rc = P->d_dinv.alloc(static_cast<size_t>(3 * (P->n + 1024)))) ||
        (rc = P->d_dkind.alloc(static_cast<size_t>(P->n + 1024))))
        return rc;
    return SMFGPU_OK;
}

// Levels of small fronts on the device, from lv0 through the run of levels
// whose symbolic fronts are all small: room for the run (an estimate that
// allows some delays) is reserved, the layout and front kernels of every
// level are launched, and the counters and the done fronts are read back
// once. A level refused for room is retried with twice the room; one
// refused for a front too large is left to the host layout (hold).
// next: the first level not done.
static int ldlt_dev_levels(smfgpu_plan_s *P, int32_t lv0, double u, double eff_small, double small,
                           const StaticPrm &S, int32_t &next, int32_t &hold) {
    hipStream_t st = P->stream;
    const int32_t nlev = static_cast<int32_t>(P->level_fronts.size());
    const int32_t symmax = std::min(kDevSymMax, P->dev_fmax);
    int32_t end = lv0;
    while (end < nlev && P->lv_fsym_max[end] <= symmax) ++end;
    next = lv0;
    hold = -1;
    int64_t grow = 1;
    int32_t retried = -1;
    int rc;
    while (next < end) {
        const double t_start = wall_s();
        const int64_t *r0 = &P->lv_room[4 * static_cast<size_t>(next)], *r1 = &P->lv_room[4 * static_cast<size_t>(end)];
        const int64_t e_sq = (r1[0] - r0[0]) * grow, e_fac = (r1[1] - r0[1]) * grow;
        const int64_t e_rows = (r1[2] - r0[2]) * grow, e_dinv = (r1[3] - r0[3]) * grow;
        const int64_t R = P->sq_pool.take(e_sq);
        if ((rc = ensure(P->d_arena, static_cast<size_t>(P->sq_pool.end), true, "front squares")) ||
            (rc = ensure(P->d_fac, static_cast<size_t>(P->fac_used + e_fac), true, "factor store")) ||
            (rc = ensure(P->d_lrows, static_cast<size_t>(P->rows_used + e_rows), true, "row ids")) ||
            (rc = ensure(P->d_ltags, static_cast<size_t>(P->rows_used + e_rows), true, "pivot tags")) ||
            (rc = ensure(P->d_dinv, static_cast<size_t>(3 * (P->dinv_used + e_dinv)), true, "D^-1 table")) ||
            (rc = ensure(P->d_dkind, static_cast<size_t>(P->dinv_used + e_dinv), true, "D^-1 kinds")))
            return rc;
        unsigned long long *h = P->h_dctr;
        h[kDcSq] = static_cast<unsigned long long>(R);
        h[kDcFac] = static_cast<unsigned long long>(P->fac_used);
        h[kDcRows] = static_cast<unsigned long long>(P->rows_used);
        h[kDcDinv] = static_cast<unsigned long long>(P->dinv_used);
        h[kDcLevel] = ~0ull;
        h[kDcWhy] = 0;
        HIP_OK(hipMemcpyAsync(P->d_dctr.p, h, kDcN * sizeof(unsigned long long), hipMemcpyHostToDevice, st));
        DevLevel L{};
        L.sym = P->d_fronts.p;
        L.sym_rows = P->d_rows.p;
        L.sym_cmap = P->d_cmap.p;
        L.sym_cmofs = P->d_cmofs.p;
        L.child_ptr = P->d_child_ptr.p;
        L.child_idx = P->d_child_idx.p;
        L.parent = P->d_parent.p;
        L.ctr = P->d_dctr.p;
        L.cap[kDcSq] = R + e_sq;
        L.cap[kDcFac] = P->fac_used + e_fac;
        L.cap[kDcRows] = P->rows_used + e_rows;
        L.cap[kDcDinv] = P->dinv_used + e_dinv;
        L.fmax = P->dev_fmax;
        const LdltDev d = ldlt_dev(P);
        for (int32_t l = next; l < end; ++l) {
            const int32_t *fl = P->d_flist.p + P->lv_flist_ofs[l];
            const int32_t nf = static_cast<int32_t>(P->lv_flist_ofs[l + 1] - P->lv_flist_ofs[l]);
            launch_dev_layout(st, d, L, fl, nf, l);
            launch_dev_fronts(st, d, L, fl, nf, l, u, eff_small, small, S, P->dev_wave != 0);
        }
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(h, P->d_dctr.p, kDcN * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        const double t_run = wall_s();
        P->t_ldlt_host[0] += t_run - t_start;
        const int32_t done = static_cast<int32_t>(std::min<unsigned long long>(h[kDcLevel], end));
        if (done > next) {
            // the whole front array (the host mirror of every front done
            // before agrees with it; later fronts are laid out again)
            HIP_OK(hipMemcpyAsync(P->fl.data(), P->d_fl.p, P->fronts.size() * sizeof(FrontLd), hipMemcpyDeviceToHost,
                                  st));
            HIP_OK(hipStreamSynchronize(st));
            int64_t fac_end = P->fac_used, rows_end = P->rows_used, dinv_end = P->dinv_used;
            if (done == end) {
                fac_end = static_cast<int64_t>(h[kDcFac]);
                rows_end = static_cast<int64_t>(h[kDcRows]);
                dinv_end = static_cast<int64_t>(h[kDcDinv]);
            } else {
                for (int32_t l = next; l < done; ++l)
                    for (int32_t s : P->level_fronts[l]) {
                        const FrontLd &F = P->fl[s];
                        fac_end = std::max(fac_end, F.lofs + static_cast<int64_t>(F.f) * F.ne);
                        rows_end = std::max(rows_end, F.rofs + F.f);
                        dinv_end = std::max(dinv_end, F.dofs + F.ne);
                    }
            }
            P->rows_h.resize(rows_end);
            HIP_OK(hipMemcpyAsync(P->rows_h.data() + P->rows_used, P->d_lrows.p + P->rows_used,
                                  (rows_end - P->rows_used) * sizeof(int32_t), hipMemcpyDeviceToHost, st));
            HIP_OK(hipStreamSynchronize(st));
            P->fac_used = fac_end;
            P->rows_used = rows_end;
            P->dinv_used = dinv_end;
            P->dev_levels += done - next;
            next = done;
        }
        P->t_ldlt_host[1] += wall_s() - t_run;
        if (next >= end) break;
        if ((h[kDcWhy] & kDevWhySize) || retried == next) {
            hold = next;
            break;
        }
        retried = next;
        grow *= 2;
    }
    return SMFGPU_OK;
}

// Large fronts of one level (multi-workgroup blocked LDL^T): the panel,
// update and stall kernels in turn until every front is done (the state
// read back every kBigCheck rounds), then the contribution blocks' update
// in chunks of an outer block's pivots and the fronts' results.
static int ldlt_big_fronts(smfgpu_plan_s *P, const LdltDev &d0, const std::vector<int32_t> &big,
                           const smfgpu_factor_options &opt, double eff_small, double small, int32_t nb,
                           std::vector<BigLd> &h, int32_t &reduced) {
    constexpr int kBigCheck = 16;
    hipStream_t st = P->stream;
    LdltDev d = d0;
    d.xpl = nullptr;
    const int32_t nbig = static_cast<int32_t>(big.size());
    const bool group = P->ldlt_group != 0 && nb == 16;
    const bool autosel = opt.precision == SMFGPU_PREC_AUTO;
    const bool collect = opt.collect != 0;
    h.assign(nbig, BigLd{});
    int64_t cap = 64;
    int32_t nmode[4] = {0, 0, 0, 0};
    const int code = autosel ? 0 : prec_code(opt.precision);
    int rc;
    // learned selector: the candidates' features from their assembled
    // squares, then (AUTO) each one's precision, read back for the launches
    std::vector<SelFront> sf;
    std::vector<int32_t> sf_big;  // index into big of each candidate
    int32_t maxne = 0, maxq = 0;
    int64_t snap_total = 0;
    if (autosel || collect) {
        int64_t cs = 0;
        for (int32_t b = 0; b < nbig; ++b) {
            const FrontLd &F = P->fl[big[b]];
            if (F.f < opt.mixed_min_front) continue;
            sf.push_back(SelFront{big[b], F.f, F.ne, F.nd, cs, F.fofs, snap_total});
            sf_big.push_back(b);
            cs += F.ne;
            snap_total += static_cast<int64_t>(F.f) * F.f;
            maxne = std::max(maxne, F.ne);
            maxq = std::max(maxq, F.f - F.ne);
        }
        if (!sf.empty()) {
            if ((rc = ensure(P->d_selfr, sf.size(), false, "selector fronts")) ||
                (rc = ensure(P->d_colstat, static_cast<size_t>(std::max<int64_t>(cs, 1)), false, "selector statistics")))
                return rc;
            HIP_OK(hipMemcpyAsync(P->d_selfr.p, sf.data(), sf.size() * sizeof(SelFront), hipMemcpyHostToDevice, st));
            launch_sq_colstat(st, P->d_arena.p, P->d_selfr.p, static_cast<int32_t>(sf.size()), maxne, P->d_colstat.p);
            launch_select(st, P->d_selfr.p, static_cast<int32_t>(sf.size()), P->d_colstat.p, sel_prm(P, opt),
                          P->d_feat.p, P->d_prec.p);
            if (autosel) {
                std::vector<int8_t> pc(P->nsup);
                HIP_OK(hipMemcpyAsync(pc.data(), P->d_prec.p, P->nsup, hipMemcpyDeviceToHost, st));
                HIP_OK(hipStreamSynchronize(st));
                for (const SelFront &x : sf) P->ldlt_prec[x.s] = pc[x.s];
            }
            if (collect)
                for (const SelFront &x : sf)
                    if (x.p < x.f) P->sel_cand.push_back(x.s);
        }
    }
    for (int32_t b = 0; b < nbig; ++b) {
        const FrontLd &F = P->fl[big[b]];
        BigLd &B = h[b];
        B.s = big[b];
        B.live = F.ne;
        B.k_paired = -1;
        B.spec = group ? 1 : 0;
        // mixed precision: reduced-precision MFMA updates on the fronts of
        // at least mixed_min_front rows (the pivot logic stays FP64), the
        // requested one or the selector's choice
        if (autosel) B.prec = F.f >= opt.mixed_min_front ? P->ldlt_prec[big[b]] : 0;
        else B.prec = code > 0 && F.f >= opt.mixed_min_front ? code : 0;
        B.outer = (P->ldlt_planes != 0 && B.prec >= 2) ? P->ldlt_outer_wide : kBigOuter;
        B.oe = std::min(F.ne, B.outer);
        P->ldlt_prec[big[b]] = static_cast<int8_t>(B.prec);
        ++nmode[B.prec];
        cap += 4 * (static_cast<int64_t>(F.ne) / std::max(nb, 1) + 2) + 2 * static_cast<int64_t>(F.ne);
    }
    reduced += nbig - nmode[0];
    if ((rc = ensure(P->d_big, static_cast<size_t>(nbig), false, "large-front states"))) return rc;
    // bf16 planes of the rank-kBigWork updates' operands (BF16 and split-BF16
    // fronts, and the training pairs' split-BF16 and BF16 runs)
    if (P->ldlt_planes != 0 && (nmode[2] + nmode[3] > 0 || (collect && !sf.empty()))) {
        int64_t planes = 0;
        for (int32_t s : big) planes = std::max(planes, 4 * (P->fl[s].wofs + static_cast<int64_t>(P->fl[s].f) * kBigWork));
        if ((rc = ensure(P->d_xpl, static_cast<size_t>(planes), false, "update planes"))) return rc;
        d.xpl = P->d_xpl.p;
    }
    HIP_OK(hipMemcpyAsync(P->d_big.p, h.data(), nbig * sizeof(BigLd), hipMemcpyHostToDevice, st));
    // update workgroups per front: a multiple of 8 (one run of tiles per XCD)
    const int32_t per_front = std::max(64, P->ldlt_update_wg / nbig) / 8 * 8;
    // speculative groups: workgroups per front over the rows below a
    // group's diagonal region, and their column maxima
    int32_t maxf = 0;
    for (int32_t s : big) maxf = std::max(maxf, P->fl[s].f);
    const int32_t nch = (maxf + kGroupRows - 1) / kGroupRows + 1;
    const size_t gstride = static_cast<size_t>(nch) * kGroupCols + kGroupBuf;
    if (group && (rc = ensure(P->d_gpart, static_cast<size_t>(nbig) * gstride, false, "group buffer"))) return rc;
    double *gpart = group ? P->d_gpart.p : nullptr;
    for (int64_t it = 0;; ++it) {
        if (group) launch_big_group_stats(st, d, P->d_big.p, nbig, gpart, nch);
        launch_big_panel(st, d, P->d_big.p, nbig, opt.pivot_u, eff_small, small, nb, gpart, nch);
        if (group) launch_big_group_apply(st, d, P->d_big.p, nbig, gpart, nch);
        launch_big_update(st, d, P->d_big.p, nbig, per_front, nmode);
        launch_big_stall(st, d, P->d_big.p, nbig, nb);
        if (it % kBigCheck == kBigCheck - 1) {
            HIP_OK(hipMemcpyAsync(h.data(), P->d_big.p, nbig * sizeof(BigLd), hipMemcpyDeviceToHost, st));
            HIP_OK(hipStreamSynchronize(st));
            bool done = true;
            for (const BigLd &B : h) done = done && B.pass == 2;
            if (done) break;
            if (it > cap) return set_error(SMFGPU_ERR_DEVICE, "factor: a large front did not finish");
        }
    }
    HIP_OK(hipGetLastError());
    int32_t pmax = 0;
    for (const BigLd &B : h) pmax = std::max(pmax, B.k);
    // contribution blocks: chunks of kBigOuter pivots, of the planes' outer
    // block for the BF16 modes on planes
    auto cb_update = [&](const int32_t *nm) {
        if (nm[0] + nm[1] > 0 || (d.xpl == nullptr && nm[2] + nm[3] > 0))
            for (int32_t c0 = 0; c0 < pmax; c0 += kBigOuter)
                launch_big_cb(st, d, P->d_big.p, nbig, c0, kBigOuter, per_front, nm, false);
        if (d.xpl != nullptr && nm[2] + nm[3] > 0)
            for (int32_t c0 = 0; c0 < pmax; c0 += P->ldlt_outer_wide)
                launch_big_cb(st, d, P->d_big.p, nbig, c0, P->ldlt_outer_wide, per_front, nm, true);
    };
    const bool labels = collect && !sf.empty() && maxq > 0;
    if (labels) {
        // training pairs: the candidates' contribution-block update run in
        // FP32, split BF16 and BF16 from the same square (copies in slots
        // 0-2 of the snapshot buffer, the square itself kept in slot 3 and
        // put back after each run), then in FP64 as the factorization's own
        if (4.0 * static_cast<double>(snap_total) * sizeof(double) > 32.0e9)
            return set_error(SMFGPU_ERR_ALLOC, "factor: training-pair snapshots would pass 32 GB");
        if ((rc = ensure(P->d_snap, static_cast<size_t>(4 * snap_total), false, "training-pair snapshots"))) return rc;
        auto copy_sq = [&](bool to_snap, int slot) -> int {
            for (const SelFront &x : sf) {
                double *sq = P->d_arena.p + x.fofs;
                double *sn = P->d_snap.p + slot * snap_total + x.sofs;
                const size_t bytes = static_cast<size_t>(x.f) * x.f * sizeof(double);
                HIP_OK(hipMemcpyAsync(to_snap ? sn : sq, to_snap ? sq : sn, bytes, hipMemcpyDeviceToDevice, st));
            }
            return SMFGPU_OK;
        };
        if ((rc = copy_sq(true, 3))) return rc;
        std::vector<int8_t> is_cand(nbig, 0);
        for (int32_t b : sf_big) is_cand[b] = 1;
        for (int m : {1, 3, 2}) {
            int32_t nm[4] = {0, 0, 0, 0};
            for (int32_t b = 0; b < nbig; ++b) {
                h[b].prec = is_cand[b] ? m : -1;  // -1: no update kernel touches the front
                if (is_cand[b]) ++nm[m];
            }
            HIP_OK(hipMemcpyAsync(P->d_big.p, h.data(), nbig * sizeof(BigLd), hipMemcpyHostToDevice, st));
            cb_update(nm);
            if ((rc = copy_sq(true, kSelOutOfCode[m])) || (rc = copy_sq(false, 3))) return rc;
        }
        for (BigLd &B : h) B.prec = 0;
        HIP_OK(hipMemcpyAsync(P->d_big.p, h.data(), nbig * sizeof(BigLd), hipMemcpyHostToDevice, st));
    }
    cb_update(nmode);
    if (labels)
        launch_label_sq(st, P->d_arena.p, P->d_snap.p, snap_total, P->d_selfr.p, static_cast<int32_t>(sf.size()), maxq,
                        P->d_lab.p);
    launch_big_finish(st, d, P->d_big.p, nbig);
    HIP_OK(hipGetLastError());
    return SMFGPU_OK;
}

// Step lists of the blocked sweeps from the realized fronts (pivots per
// front known after the factorization): 64-pivot blocks; forward steps
// solve block t after block t-1's update of its rows, while chunk items
// apply that update to the rest of the front (64-row chunks, rows past the
// pivots by atomic add); backward steps first take the contributions of
// the rows past the pivots (512-row segments, atomic), then solve the
// blocks last to first, chunk items applying each solved block to the
// columns before it. One launch per step covers every front of a level.
// The items of one front's forward (bwd = false) or backward sweep, step
// by step: emit(step, item).
template <class Emit>
static void front_solve_items(int32_t f, int32_t p, int32_t s, bool bwd, Emit &&emit) {
    // B pivots per step (the kernels solve them as 64-pivot sub-blocks),
    // chunks of C rows (forward) or columns (backward)
    constexpr int32_t B = 256, C = 64, Seg = 512, Tri = 2;
    if (p <= 0) return;
    const int32_t nb = (p + B - 1) / B;
    if (!bwd) {
        constexpr int32_t Fused = 4, FusedRows = 512;
        if (nb == 1 && f > p && f - p <= FusedRows) {
            // one workgroup: the triangle, then the rows below it
            emit(0, SolveItem{s, 0, 0, p, f, 0, p, Tri | Fused});
            return;
        }
        // block b: a step of row items applying block b - 1 to every row
        // from block b down, then a step of its triangle item alone (so the
        // triangle's one workgroup does no more than the block's own work)
        int32_t t = 0;
        emit(t++, SolveItem{s, 0, 0, 0, 0, 0, std::min(B, p), Tri});
        for (int32_t b = 1; b < nb; ++b, ++t) {
            const int32_t tk = b * B, tkb = std::min(B, p - tk), k = tk - B;
            for (int32_t r = tk; r < f; r += C) emit(t, SolveItem{s, k, B, r, std::min(r + C, f), 0, 0, 0});
            emit(++t, SolveItem{s, 0, 0, 0, 0, tk, tkb, Tri});
        }
        if (f > p) {
            const int32_t k = (nb - 1) * B;
            for (int32_t r = p; r < f; r += C) emit(t, SolveItem{s, k, p - k, r, std::min(r + C, f), 0, 0, 0});
        }
        return;
    }
    if (nb == 1 && f > p && f - p <= Seg) {
        // one workgroup: the rows below (one segment) folded into the
        // triangle item's own column sums, then the triangle
        emit(0, SolveItem{s, p, f - p, 0, 0, 0, p, Tri});
        return;
    }
    int32_t t = 0;
    if (f > p) {
        for (int32_t r = p; r < f; r += Seg)
            for (int32_t c = 0; c < p; c += C) emit(0, SolveItem{s, r, std::min(Seg, f - r), c, std::min(c + C, p), 0, 0, 1});
        t = 1;
    }
    const int32_t kL = (nb - 1) * B;
    emit(t++, SolveItem{s, 0, 0, 0, 0, kL, p - kL, Tri});
    // block j: a step of column items applying block j + 1 to every column
    // up to block j's last, then a step of its triangle item alone
    for (int32_t j = nb - 2; j >= 0; --j, ++t) {
        const int32_t tk = j * B, k = tk + B, kb = std::min(B, p - k);
        for (int32_t c = 0; c < tk + B; c += C) emit(t, SolveItem{s, k, kb, c, c + C, 0, 0, 0});
        emit(++t, SolveItem{s, 0, 0, 0, 0, tk, B, Tri});
    }
}

// Solve schedule of the factorization: per level, the items of each step
// (items of all the level's fronts), counted first and then written in
// place.
static int ldlt_solve_lists(smfgpu_plan_s *P) {
    const int32_t nlev = static_cast<int32_t>(P->level_fronts.size());
    P->fwd_steps.clear();
    P->bwd_steps.clear();
    P->fwd_lv.assign(nlev + 1, 0);
    P->bwd_lv.assign(nlev + 1, 0);
    std::vector<int64_t> cnt, pos;
    // (rows, stored pivots) of every front: one sequential pass over the
    // front array; the passes below read these compact arrays
    const size_t ns = P->fronts.size();
    std::vector<int32_t> lf(ns), lp(ns);
    for (size_t s = 0; s < ns; ++s) {
        lf[s] = P->fl[s].f;
        lp[s] = P->fl[s].pstore;
    }
    // counts per level and step give the steps' offsets; the items are then
    // written in place into pinned staging and uploaded once
    auto build = [&](bool bwd, PinnedArr<SolveItem> &items, DBuf<SolveItem> &dev,
                     std::vector<std::pair<int64_t, int64_t>> &steps, std::vector<int32_t> &lvs) -> int {
        std::vector<int32_t> lv_nstep(nlev, 0);
        std::vector<int64_t> lv_t0(nlev, 0), tfirst;
        int64_t total = 0;
        for (int32_t lv = 0; lv < nlev; ++lv) {
            lvs[lv] = static_cast<int32_t>(steps.size());
            cnt.clear();
            for (int32_t s : P->level_fronts[lv])
                front_solve_items(lf[s], lp[s], s, bwd, [&](int32_t t, const SolveItem &) {
                    if (cnt.size() <= static_cast<size_t>(t)) cnt.resize(t + 1, 0);
                    ++cnt[t];
                });
            lv_t0[lv] = static_cast<int64_t>(tfirst.size());
            for (size_t t = 0; t < cnt.size(); ++t) {
                tfirst.push_back(total);  // first item of step index t (none if cnt[t] == 0)
                if (cnt[t] > 0) {
                    steps.emplace_back(total, cnt[t]);
                    total += cnt[t];
                }
            }
            lv_nstep[lv] = static_cast<int32_t>(cnt.size());
        }
        lvs[nlev] = static_cast<int32_t>(steps.size());
        int rc;
        if ((rc = items.reserve(static_cast<size_t>(std::max<int64_t>(total, 1))))) return rc;
        for (int32_t lv = 0; lv < nlev; ++lv) {
            // the first item of each of the level's step indices, from the
            // counting pass
            pos.assign(tfirst.begin() + lv_t0[lv], tfirst.begin() + lv_t0[lv] + lv_nstep[lv]);
            for (int32_t s : P->level_fronts[lv])
                front_solve_items(lf[s], lp[s], s, bwd, [&](int32_t t, const SolveItem &it) { items[pos[t]++] = it; });
        }
        for (const auto &st : steps)
            if (st.second > INT32_MAX) return set_error(SMFGPU_ERR_ALLOC, "factor: solve step too large");
        if ((rc = ensure(dev, static_cast<size_t>(std::max<int64_t>(total, 1)), false,
                         bwd ? "backward solve items" : "forward solve items")))
            return rc;
        if (total > 0)
            HIP_OK(hipMemcpyAsync(dev.p, items.data(), static_cast<size_t>(total) * sizeof(SolveItem),
                                  hipMemcpyHostToDevice, P->stream));
        return SMFGPU_OK;
    };
    int rc;
    if ((rc = build(false, P->h_fwd_items, P->d_fwd_items, P->fwd_steps, P->fwd_lv)) ||
        (rc = build(true, P->h_bwd_items, P->d_bwd_items, P->bwd_steps, P->bwd_lv)))
        return rc;
    HIP_OK(hipStreamSynchronize(P->stream));
    return SMFGPU_OK;
}

int ldlt_factor(smfgpu_plan_s *P, const double *dv, int64_t nvalues, const smfgpu_factor_options &opt,
                smfgpu_factor_info &fi) {
    hipStream_t st = P->stream;
    const int32_t n = P->n;
    P->ldlt_ready = false;
    launch_map_values(st, dv, nvalues, P->d_slot_ptr.p, P->d_slot_src.p, P->nnz_perm, P->d_av.p, P->d_red.p);
    // S A S, max|S A S| in place of max|A| (factor_indef.cpp:508-559, 608-613)
    if (P->scaled)
        launch_scale_values(st, n, P->d_fr_ptr.p, P->d_fr_col.p, P->d_fr_slot.p, P->d_perm.p, P->d_scale.p, P->d_av.p,
                            P->d_av_raw.p, P->d_red.p);
    unsigned long long maxbits = 0;
    HIP_OK(hipMemcpyAsync(&maxbits, P->d_red.p, sizeof(maxbits), hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    double max_abs;
    std::memcpy(&max_abs, &maxbits, sizeof(double));
    fi.max_abs = max_abs;
    if (!(maxbits < 0x7ff0000000000000ull)) {
        fi.status = SMFGPU_STATUS_SINGULAR;
        return set_error(SMFGPU_ERR_VALUE, "factor: non-finite input value");
    }
    // eff_small (factor_indef.cpp:633-636)
    const double small = opt.small_pivot;
    const double eff_small = (max_abs > 0.0 && max_abs < small) ? small * max_abs : small;
    const int32_t nb = std::max(2, opt.panel);
    // static pivoting values (factor_indef.cpp:642-645)
    StaticPrm spr{};
    spr.sp = opt.static_pivot_tol > 0.0 ? opt.static_pivot_tol * max_abs : 0.0;
    spr.fallback = opt.fallback_static_pivot_tol > 0.0 ? opt.fallback_static_pivot_tol * max_abs : 0.0;
    spr.flags = opt.static_flags;
    P->fac_used = P->rows_used = P->dinv_used = 0;
    P->sq_pool.reset();
    P->ldlt_prec.assign(P->nsup, 0);
    P->sel_cand.clear();
    std::vector<int32_t> pos(n, -1);
    std::vector<int32_t> cmap;
    std::vector<AsmTask> tasks;
    std::vector<ChildCol> cc;
    std::vector<int32_t> colstart, colfill;
    std::vector<int32_t> ids, small_list, blocked_list, big_list, ep_list, store_list;
    std::vector<BigLd> big_h;
    std::vector<FrontLd> patch;
    std::vector<int32_t> lrows;
    const int32_t nlev = static_cast<int32_t>(P->level_fronts.size());
    int rc;
    int32_t levels_done = 0, reduced = 0;
    bool inertia_exceeded = false, work_exceeded = false;
    // Early stops (multifrontal_serial.cpp:653-671, 1196-1205, 1612-1622).
    // The reference takes the fronts one by one in supernode order (its
    // postorder is the identity, symbolic_analysis.cpp:1200-1237): before a
    // front, a work limit stops the run when the front's work num_elig * f^2
    // would take the total past floor(work_limit) (that work is counted, the
    // front is not factored); after it, max_negative_pivots stops the run
    // once the negative pivots exceed it. Levels here finish fronts out of
    // that order, so a cursor walks the supernodes in order through the
    // finished ones and applies both tests there: the stopping front and the
    // partial counts (the fronts before it, and for the inertia stop the
    // front itself) are the reference's. A front's results do not depend on
    // when it runs, only on its children.
    const bool work_armed = opt.work_limit_on != 0 && opt.work_limit >= 0.0;
    const bool neg_armed = opt.max_negative_pivots >= 0;
    const int64_t work_cap = !work_armed ? -1
                             : (opt.work_limit >= 9.0e18 ? INT64_MAX
                                                         : static_cast<int64_t>(std::floor(opt.work_limit)));
    std::vector<int64_t> fwork;  // per supernode once laid out, -1 before
    std::vector<char> fdone;     // per supernode once factored (armed runs)
    if (work_armed || neg_armed) {
        fwork.assign(P->nsup, -1);
        fdone.assign(P->nsup, 0);
    }
    int32_t cur = 0, stop_s = -1;
    int64_t work_prefix = 0, neg_prefix = 0;
    P->rows_h.resize(0);
    P->dev_levels = 0;
    for (double &t : P->t_ldlt_host) t = 0.0;
    // levels of small fronts run on the device unless a limit needs the
    // per-level checks
    const bool dev_on = P->dev_fmax > 0 && !work_armed && opt.max_negative_pivots < 0;
    int32_t dev_hold = -1;
    for (int32_t lv = 0; lv < nlev; ++lv) {
        if (dev_on && lv != dev_hold && P->lv_fsym_max[lv] <= std::min(kDevSymMax, P->dev_fmax)) {
            int32_t next = lv, hold = -1;
            if ((rc = ldlt_dev_levels(P, lv, opt.pivot_u, eff_small, small, spr, next, hold))) return rc;
            levels_done = next;
            dev_hold = hold;
            if (next > lv) {
                lv = next - 1;
                continue;
            }
            dev_hold = lv;
        }
        const auto &fr = P->level_fronts[lv];
        cmap.clear();
        tasks.clear();
        cc.clear();
        ids.clear();
        patch.clear();
        lrows.clear();
        small_list.clear();
        blocked_list.clear();
        big_list.clear();
        ep_list.clear();
        store_list.clear();
        const int64_t rows_start = P->rows_used;
        int64_t work_used = 0;
        for (int32_t s : fr) {
            const FrontDev &Sy = P->fronts[s];
            FrontLd F{};
            int32_t nd = 0;
            for (int32_t c = P->child_ptr[s]; c < P->child_ptr[s + 1]; ++c) {
                const FrontLd &C = P->fl[P->child_idx[c]];
                // a child passes at most its eligible columns up
                if (C.delayed < 0 || C.delayed > C.ne || C.f > n)
                    return set_error(SMFGPU_ERR_DEVICE, "factor: front " + std::to_string(P->child_idx[c]) +
                                                            " came back inconsistent (f " + std::to_string(C.f) +
                                                            ", eligible " + std::to_string(C.ne) + ", delayed " +
                                                            std::to_string(C.delayed) + ", level " +
                                                            std::to_string(lv) + ")");
                nd += C.delayed;
            }
            F.nd = nd;
            F.psym = Sy.p;
            F.col0 = Sy.col0;
            F.f = nd + Sy.f;
            F.ne = nd + Sy.p;
            F.root = P->sn_parent[s] < 0 ? 1 : 0;
            F.fofs = P->sq_pool.take(static_cast<int64_t>(F.f) * F.f);
            F.lofs = P->fac_used;
            F.rofs = P->rows_used;
            F.dofs = P->dinv_used;
            F.wofs = 0;
            // the large-front state machine has no static path: a static
            // front takes the one-workgroup blocked kernel
            const bool stat = front_static_value(spr, F.root, F.nd, F.psym) > 0.0;
            if (F.f >= P->ldlt_big_min && F.ne > 0 && !stat) {
                F.wofs = work_used;
                work_used += static_cast<int64_t>(F.f) * kBigWork;
                big_list.push_back(s);
            } else if (F.f >= kBlockedMinFront) {
                F.wofs = work_used;
                work_used += static_cast<int64_t>(F.f) * nb;
                blocked_list.push_back(s);
            } else {
                small_list.push_back(s);
            }
            if (!fwork.empty()) fwork[s] = static_cast<int64_t>(F.ne) * F.f * F.f;
            P->fac_used += static_cast<int64_t>(F.f) * F.ne;
            // large stored blocks get the many-workgroup store kernel
            if (static_cast<int64_t>(F.f) * F.ne >= kStoreWide) store_list.push_back(s);
            else ep_list.push_back(s);
            P->rows_used += F.f;
            P->dinv_used += F.ne;
            // rows: children's delayed columns in child order, then the
            // symbolic rows
            int32_t at = 0;
            const size_t r0 = lrows.size();
            for (int32_t c = P->child_ptr[s]; c < P->child_ptr[s + 1]; ++c) {
                const FrontLd &C = P->fl[P->child_idx[c]];
                const int32_t *cr = P->rows_h.data() + C.rofs;
                for (int32_t t = 0; t < C.delayed; ++t) {
                    lrows.push_back(cr[C.p + t]);
                    pos[cr[C.p + t]] = at++;
                }
            }
            for (int32_t t = P->front_row_ptr[s]; t < P->front_row_ptr[s + 1]; ++t) {
                lrows.push_back(P->front_row_idx[t]);
                pos[P->front_row_idx[t]] = at++;
            }
            // child maps (increasing: delayed rows come first in both) and
            // the contributions per parent column
            const size_t cm0 = cmap.size();
            for (int32_t c = P->child_ptr[s]; c < P->child_ptr[s + 1]; ++c) {
                const int32_t ch = P->child_idx[c];
                FrontLd &C = P->fl[ch];
                const int32_t q = C.f - C.p;
                C.cmofs = static_cast<int64_t>(cmap.size());
                const int32_t *cr = P->rows_h.data() + C.rofs + C.p;
                int32_t prev = -1;
                for (int32_t ii = 0; ii < q; ++ii) {
                    const int32_t l = pos[cr[ii]];
                    if (l < 0 || l <= prev)
                        return set_error(SMFGPU_ERR_INVALID_ARG, "factor: child row outside its parent front");
                    cmap.push_back(l);
                    prev = l;
                }
                ids.push_back(ch);
                patch.push_back(C);
            }
            // per parent column, its children's columns in child order (a
            // counting sort over the maps just built)
            colstart.assign(static_cast<size_t>(F.f) + 1, 0);
            for (size_t e = cm0; e < cmap.size(); ++e) ++colstart[cmap[e] + 1];
            for (int32_t c = 0; c < F.f; ++c) colstart[c + 1] += colstart[c];
            const int32_t cc0 = static_cast<int32_t>(cc.size());
            cc.resize(cc.size() + (cmap.size() - cm0));
            colfill.assign(colstart.begin(), colstart.end() - 1);
            for (int32_t c = P->child_ptr[s]; c < P->child_ptr[s + 1]; ++c) {
                const int32_t ch = P->child_idx[c];
                const FrontLd &C = P->fl[ch];
                const int32_t q = C.f - C.p;
                for (int32_t jj = 0; jj < q; ++jj) cc[cc0 + colfill[cmap[C.cmofs + jj]]++] = ChildCol{ch, jj};
            }
            for (int32_t c = 0; c < F.f; ++c) tasks.push_back(AsmTask{s, c, cc0 + colstart[c], cc0 + colstart[c + 1]});
            for (size_t t = r0; t < lrows.size(); ++t) pos[lrows[t]] = -1;
            P->fl[s] = F;
            ids.push_back(s);
            patch.push_back(F);
        }
        // the next front in order is in this level and its work crosses the
        // limit: stop before it, with nothing of this level launched
        if (work_armed && cur < P->nsup && !fdone[cur] && fwork[cur] >= 0 && work_prefix + fwork[cur] > work_cap) {
            work_exceeded = true;
            stop_s = cur;
            work_prefix += fwork[cur];
            break;
        }
        if (cc.size() > static_cast<size_t>(INT32_MAX)) return set_error(SMFGPU_ERR_ALLOC, "factor: level too large");
        // capacities (arena contents stay: offsets are relative)
        if ((rc = ensure(P->d_arena, static_cast<size_t>(P->sq_pool.end), true, "front squares")) ||
            (rc = ensure(P->d_fac, static_cast<size_t>(P->fac_used), true, "factor store")) ||
            (rc = ensure(P->d_lrows, static_cast<size_t>(P->rows_used), true, "row ids")) ||
            (rc = ensure(P->d_ltags, static_cast<size_t>(P->rows_used), true, "pivot tags")) ||
            (rc = ensure(P->d_dinv, static_cast<size_t>(3 * P->dinv_used), true, "D^-1 table")) ||
            (rc = ensure(P->d_dkind, static_cast<size_t>(P->dinv_used), true, "D^-1 kinds")) ||
            (rc = ensure(P->d_work, static_cast<size_t>(std::max<int64_t>(work_used, 1)), false, "panel workspace")) ||
            (rc = ensure(P->d_lcmap, std::max<size_t>(cmap.size(), 1), false, "child maps")))
            return rc;
        // one staged upload: patch, ids, rows, cmap, tasks, contributions,
        // front lists
        P->h_stage.clear();
        const size_t o_patch = stage(P->h_stage, patch);
        const size_t o_ids = stage(P->h_stage, ids);
        const size_t o_rows = stage(P->h_stage, lrows);
        const size_t o_cmap = stage(P->h_stage, cmap);
        const size_t o_tasks = stage(P->h_stage, tasks);
        const size_t o_cc = stage(P->h_stage, cc);
        const size_t o_small = stage(P->h_stage, small_list);
        const size_t o_blk = stage(P->h_stage, blocked_list);
        const size_t o_fr = stage(P->h_stage, fr);
        const int32_t n_inline = static_cast<int32_t>(ep_list.size());
        int32_t max_store_cols = 0;
        for (int32_t s : store_list) max_store_cols = std::max(max_store_cols, P->fl[s].ne);
        ep_list.insert(ep_list.end(), store_list.begin(), store_list.end());
        const size_t o_ep = stage(P->h_stage, ep_list);
        const size_t o_back = (P->h_stage.size() + 7) & ~size_t{7};
        P->h_stage.resize(o_back + fr.size() * sizeof(FrontLd));
        if ((rc = ensure(P->d_stage, P->h_stage.size(), false, "level staging"))) return rc;
        unsigned char *S = P->d_stage.p;
        HIP_OK(hipMemcpyAsync(S, P->h_stage.data(), o_back, hipMemcpyHostToDevice, st));
        HIP_OK(hipMemcpyAsync(P->d_lrows.p + rows_start, S + o_rows, lrows.size() * sizeof(int32_t),
                              hipMemcpyDeviceToDevice, st));
        HIP_OK(hipMemsetAsync(P->d_ltags.p + rows_start, 0, lrows.size(), st));
        if (!cmap.empty())
            HIP_OK(hipMemcpyAsync(P->d_lcmap.p, S + o_cmap, cmap.size() * sizeof(int32_t), hipMemcpyDeviceToDevice, st));
        launch_front_patch(st, P->d_fl.p, reinterpret_cast<const FrontLd *>(S + o_patch),
                           reinterpret_cast<const int32_t *>(S + o_ids), static_cast<int32_t>(ids.size()));
        LdltDev d = ldlt_dev(P);
        launch_ldlt_assemble(st, d, reinterpret_cast<const AsmTask *>(S + o_tasks), static_cast<int64_t>(tasks.size()),
                             reinterpret_cast<const ChildCol *>(S + o_cc));
        // the children's squares are free for the launches that follow
        for (int32_t s : fr)
            for (int32_t c = P->child_ptr[s]; c < P->child_ptr[s + 1]; ++c) {
                const FrontLd &C = P->fl[P->child_idx[c]];
                if (!C.dev) P->sq_pool.give(C.fofs, static_cast<int64_t>(C.f) * C.f);
            }
        launch_ldlt_small(st, d, reinterpret_cast<const int32_t *>(S + o_small), static_cast<int32_t>(small_list.size()),
                          opt.pivot_u, eff_small, small, spr);
        launch_ldlt_blocked(st, d, reinterpret_cast<const int32_t *>(S + o_blk),
                            static_cast<int32_t>(blocked_list.size()), opt.pivot_u, eff_small, small, nb, spr);
        if (!big_list.empty() && (rc = ldlt_big_fronts(P, d, big_list, opt, eff_small, small, nb, big_h, reduced)))
            return rc;
        launch_ldlt_epilogue(st, d, reinterpret_cast<const int32_t *>(S + o_ep), static_cast<int32_t>(ep_list.size()),
                             n_inline);
        launch_ldlt_store(st, d, reinterpret_cast<const int32_t *>(S + o_ep) + n_inline,
                          static_cast<int32_t>(store_list.size()), max_store_cols);
        for (int32_t s : fr)
            if (P->fl[s].root) P->sq_pool.give(P->fl[s].fofs, static_cast<int64_t>(P->fl[s].f) * P->fl[s].f);
        launch_front_gather(st, P->d_fl.p, reinterpret_cast<FrontLd *>(S + o_back),
                            reinterpret_cast<const int32_t *>(S + o_fr), static_cast<int32_t>(fr.size()));
        HIP_OK(hipGetLastError());
        // read back the level's fronts and final row orders
        std::vector<FrontLd> back(fr.size());
        HIP_OK(hipMemcpyAsync(back.data(), S + o_back, back.size() * sizeof(FrontLd), hipMemcpyDeviceToHost, st));
        P->rows_h.resize(P->rows_used);
        HIP_OK(hipMemcpyAsync(P->rows_h.data() + rows_start, P->d_lrows.p + rows_start, lrows.size() * sizeof(int32_t),
                              hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        for (size_t i = 0; i < fr.size(); ++i) P->fl[fr[i]] = back[i];
        levels_done = lv + 1;
        if (!fdone.empty()) {
            for (int32_t s : fr) fdone[s] = 1;
            while (cur < P->nsup && fdone[cur]) {
                if (work_armed && work_prefix + fwork[cur] > work_cap) {
                    work_exceeded = true;
                    work_prefix += fwork[cur];
                    break;
                }
                work_prefix += fwork[cur];
                neg_prefix += P->fl[cur].nneg;
                if (neg_armed && neg_prefix > opt.max_negative_pivots) {
                    inertia_exceeded = true;
                    break;
                }
                ++cur;
            }
            if (work_exceeded || inertia_exceeded) {
                stop_s = cur;
                break;
            }
        }
    }
    const double t_stats = wall_s();
    // statistics (Info fields, factor_indef.cpp:725-770) over the fronts
    // factored
    fi.num_positive = fi.num_negative = fi.num_zero = 0;
    fi.delayed_pivots = fi.two_by_two_pivots = 0;
    fi.actual_flops = 0;
    fi.front_work = 0;
    fi.max_front = 0;
    fi.static_pivots = 0;
    fi.pivot_growth = 0.0;
    int64_t fent = 0;
    // every front, or those the reference factors before its stop (the
    // stopping front too after an inertia stop), in supernode order
    std::vector<int32_t> done_fronts;
    const int32_t n_done = stop_s < 0 ? P->nsup : (inertia_exceeded ? stop_s + 1 : stop_s);
    done_fronts.resize(n_done);
    for (int32_t s = 0; s < n_done; ++s) done_fronts[s] = s;
    for (int32_t s : done_fronts) {
        const FrontLd &F = P->fl[s];
        fi.num_positive += F.npos;
        fi.num_negative += F.nneg;
        fi.num_zero += F.nzero;
        fi.delayed_pivots += F.delayed;
        fi.two_by_two_pivots += F.n2x2;
        fi.actual_flops += F.flops;
        fi.front_work += static_cast<double>(F.ne) * F.f * static_cast<double>(F.f);
        fi.max_front = std::max(fi.max_front, F.f);
        fi.static_pivots += F.nstatic;
        fi.pivot_growth = std::max(fi.pivot_growth, F.growth);
        fent += static_cast<int64_t>(F.f) * F.pstore;
    }
    fi.factor_entries = fent;
    fi.fronts_reduced = reduced;
    if (inertia_exceeded) {
        fi.status = SMFGPU_STATUS_INERTIA_EXCEEDED;  // factor_indef.cpp:682-715: unusable
        return SMFGPU_OK;
    }
    if (work_exceeded) {
        // factor_indef.cpp:655-680: stopped at the scale-on-delay threshold,
        // the factor is unusable; the front work counts the stopping front
        fi.status = SMFGPU_STATUS_MAX_PIVOT_DELAYS;
        fi.front_work = static_cast<double>(work_prefix);
        return SMFGPU_OK;
    }
    fi.status = fi.num_zero > 0 ? SMFGPU_STATUS_SINGULAR : SMFGPU_STATUS_SUCCESS;
    const double t_lists = wall_s();
    P->t_ldlt_host[2] = t_lists - t_stats;
    if ((rc = ldlt_solve_lists(P))) return rc;
    P->t_ldlt_host[3] = wall_s() - t_lists;
    P->ldlt_ready = true;
    return SMFGPU_OK;
}

int ldlt_solve(smfgpu_plan_s *P, const double *b_dev, int64_t ldb, double *out_dev, int64_t ldo, int32_t nrhs,
               int32_t job, double *ybuf, int64_t ycap) {
    hipStream_t st = P->stream;
    const int32_t n = P->n;
    LdltDev d = ldlt_dev(P);
    const int32_t nl = static_cast<int32_t>(P->level_fronts.size());
    const bool tm = P->timing && job == 0 && !ybuf;
    // right-hand sides per pass through the sweeps: 16, or what a
    // caller-given work block holds
    const int32_t group = ybuf ? static_cast<int32_t>(std::min<int64_t>(16, ycap / std::max(n, 1))) : 16;
    if (group < 1) return set_error(SMFGPU_ERR_INVALID_ARG, "solve: work block smaller than one column");
    if (tm && P->sev.size() < 3) {
        for (size_t i = P->sev.size(); i < 3; ++i) {
            hipEvent_t e;
            HIP_OK(hipEventCreate(&e));
            P->sev.push_back(e);
        }
    }
    for (int32_t r0 = 0; r0 < nrhs; r0 += group) {
        const int32_t nr = std::min(group, nrhs - r0);
        if (!ybuf && P->d_y.n < static_cast<size_t>(n) * nr) {
            const int rc = P->d_y.alloc(static_cast<size_t>(n) * 16);
            if (rc) return rc;
        }
        double *y = ybuf ? ybuf : P->d_y.p;
        const bool fwd = job == 0 || job == 1;
        const double *sc = P->scaled ? P->d_scale.p : nullptr;
        if (fwd) launch_gather_rhs(st, P->d_perm.p, b_dev + r0 * ldb, ldb, y, n, nr, sc);
        else
            HIP_OK(hipMemcpy2DAsync(y, n * sizeof(double), b_dev + r0 * ldb, ldb * sizeof(double),
                                    n * sizeof(double), nr, hipMemcpyDeviceToDevice, st));
        if (tm) HIP_OK(hipEventRecord(P->sev[0], st));
        if (fwd)
            for (int32_t lv = 0; lv < nl; ++lv)
                for (int32_t i = P->fwd_lv[lv]; i < P->fwd_lv[lv + 1]; ++i)
                    launch_ldlt_fwd_step(st, d, P->d_fwd_items.p + P->fwd_steps[i].first, P->fwd_steps[i].second,
                                         y, n, nr);
        if (job == 2) launch_ldlt_diag_div(st, d, P->d_flist.p, static_cast<int32_t>(P->lv_flist_ofs[nl]), y, n, nr);
        if (job == 1 || job == 2) {
            // the permuted-order result
            HIP_OK(hipMemcpy2DAsync(out_dev + r0 * ldo, ldo * sizeof(double), y, n * sizeof(double),
                                    n * sizeof(double), nr, hipMemcpyDeviceToDevice, st));
            continue;
        }
        if (job == 0 || job == 4)
            launch_ldlt_diag(st, d, P->d_flist.p, static_cast<int32_t>(P->lv_flist_ofs[nl]), y, n, nr);
        if (tm) HIP_OK(hipEventRecord(P->sev[1], st));
        for (int32_t lv = nl - 1; lv >= 0; --lv)
            for (int32_t i = P->bwd_lv[lv]; i < P->bwd_lv[lv + 1]; ++i)
                launch_ldlt_bwd_step(st, d, P->d_bwd_items.p + P->bwd_steps[i].first, P->bwd_steps[i].second,
                                     y, n, nr);
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

// FactorKeep layout: steps in supernode order with the offsets of a
// serial factorization (row_offset += f, factor_offset += f * pstore).
int ldlt_export(smfgpu_plan_s *P, int64_t *sizes, int64_t *steps, double *fv, int32_t *rows, int8_t *tags) {
    const int32_t ns = P->nsup;
    int64_t nf = 0, nr = 0, nt = 0;
    for (int32_t s = 0; s < ns; ++s) {
        nf += static_cast<int64_t>(P->fl[s].f) * P->fl[s].pstore;
        nr += P->fl[s].f;
        nt += P->fl[s].pstore;
    }
    if (sizes) {
        sizes[0] = ns;
        sizes[1] = nf;
        sizes[2] = nr;
        sizes[3] = nt;
    }
    int64_t fo = 0, ro = 0, to = 0;
    std::vector<int8_t> tg;
    for (int32_t s = 0; s < ns; ++s) {
        const FrontLd &F = P->fl[s];
        if (steps) {
            steps[4 * s + 0] = ro;
            steps[4 * s + 1] = fo;
            steps[4 * s + 2] = F.pstore;
            steps[4 * s + 3] = F.f;
        }
        if (fv && F.pstore > 0)
            HIP_OK(hipMemcpy(fv + fo, P->d_fac.p + F.lofs, static_cast<size_t>(F.f) * F.pstore * sizeof(double),
                             hipMemcpyDeviceToHost));
        if (rows) std::memcpy(rows + ro, P->rows_h.data() + F.rofs, static_cast<size_t>(F.f) * sizeof(int32_t));
        if (tags && F.pstore > 0) HIP_OK(hipMemcpy(tags + to, P->d_ltags.p + F.rofs, F.pstore, hipMemcpyDeviceToHost));
        fo += static_cast<int64_t>(F.f) * F.pstore;
        ro += F.f;
        to += F.pstore;
    }
    return SMFGPU_OK;
}

}  // namespace smfgpu
