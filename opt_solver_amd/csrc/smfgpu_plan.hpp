// Host-side state of a device plan and of a factorization (shared by the
// SPD and LDL^T drivers).
#pragma once

#include "smfgpu_internal.hpp"
#include "smfgpu_ldlt.hpp"

#include <climits>
#include <map>
#include <set>
#include <string>
#include <vector>

namespace smfgpu {

#define HIP_OK(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return set_error(SMFGPU_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Device buffer with RAII.
template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    int alloc(size_t count) {
        release();
        if (count == 0) return SMFGPU_OK;
        if (hipMalloc(reinterpret_cast<void **>(&p), count * sizeof(T)) != hipSuccess) {
            p = nullptr;
            return set_error(SMFGPU_ERR_ALLOC, "device allocation of " + std::to_string(count * sizeof(T)) + " bytes failed");
        }
        n = count;
        return SMFGPU_OK;
    }
    int upload(const std::vector<T> &v) {
        int rc = alloc(v.size());
        if (rc != SMFGPU_OK) return rc;
        if (!v.empty() && hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
            return set_error(SMFGPU_ERR_DEVICE, "upload failed");
        return SMFGPU_OK;
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Host array in pinned memory: device-to-host copies land in it directly.
template <class T>
struct PinnedArr {
    T *p = nullptr;
    size_t n = 0;
    PinnedArr() = default;
    PinnedArr(const PinnedArr &) = delete;
    PinnedArr &operator=(const PinnedArr &) = delete;
    ~PinnedArr() {
        if (p) (void)hipHostFree(p);
    }
    int assign(size_t count, const T &v) {
        if (count > n) {
            if (p) (void)hipHostFree(p);
            p = nullptr;
            n = 0;
            if (hipHostMalloc(reinterpret_cast<void **>(&p), count * sizeof(T), hipHostMallocDefault) != hipSuccess) {
                p = nullptr;
                return set_error(SMFGPU_ERR_ALLOC, "pinned host allocation of " + std::to_string(count * sizeof(T)) +
                                                       " bytes failed");
            }
            n = count;
        }
        for (size_t i = 0; i < count; ++i) p[i] = v;
        return SMFGPU_OK;
    }
    // room for count elements; the contents are not kept
    int reserve(size_t count) {
        if (count <= n) return SMFGPU_OK;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
        if (hipHostMalloc(reinterpret_cast<void **>(&p), count * sizeof(T), hipHostMallocDefault) != hipSuccess) {
            p = nullptr;
            return set_error(SMFGPU_ERR_ALLOC, "pinned host allocation of " + std::to_string(count * sizeof(T)) +
                                                   " bytes failed");
        }
        n = count;
        return SMFGPU_OK;
    }
    T &operator[](size_t i) { return p[i]; }
    const T &operator[](size_t i) const { return p[i]; }
    T *data() { return p; }
};

// Host-side best-fit allocator of offsets (in elements) into one device
// buffer; freed ranges merge with free neighbours. The LDL^T driver takes
// each front's square from it when the front's level is laid out and gives
// it back once the parent has assembled the contribution block, so the
// buffer holds the live squares only.
struct RangePool {
    static constexpr int64_t kAlign = 64;
    std::map<int64_t, int64_t> by_off;               // free ranges: offset -> length
    std::multimap<int64_t, int64_t> by_len;          // length -> offset
    int64_t end = 0;                                 // high-water mark
    void reset() {
        by_off.clear();
        by_len.clear();
        end = 0;
    }
    static int64_t round(int64_t len) { return (len + kAlign - 1) / kAlign * kAlign; }
    void erase_len(int64_t len, int64_t off) {
        auto r = by_len.equal_range(len);
        for (auto it = r.first; it != r.second; ++it)
            if (it->second == off) {
                by_len.erase(it);
                return;
            }
    }
    int64_t take(int64_t len) {
        len = round(len);
        auto it = by_len.lower_bound(len);
        if (it == by_len.end()) {
            // the last free range may touch the end: extend it
            if (!by_off.empty()) {
                auto last = std::prev(by_off.end());
                if (last->first + last->second == end) {
                    const int64_t off = last->first;
                    erase_len(last->second, off);
                    by_off.erase(last);
                    end = off + len;
                    return off;
                }
            }
            const int64_t off = end;
            end += len;
            return off;
        }
        const int64_t flen = it->first, off = it->second;
        by_len.erase(it);
        by_off.erase(off);
        if (flen > len) {
            by_off[off + len] = flen - len;
            by_len.emplace(flen - len, off + len);
        }
        return off;
    }
    void give(int64_t off, int64_t len) {
        len = round(len);
        auto nx = by_off.lower_bound(off);
        if (nx != by_off.end() && off + len == nx->first) {
            len += nx->second;
            erase_len(nx->second, nx->first);
            nx = by_off.erase(nx);
        }
        if (nx != by_off.begin()) {
            auto pv = std::prev(nx);
            if (pv->first + pv->second == off) {
                off = pv->first;
                len += pv->second;
                erase_len(pv->second, pv->first);
                by_off.erase(pv);
            }
        }
        by_off[off] = len;
        by_len.emplace(len, off);
    }
};

constexpr int kPanelOuter = 512;  // default outer panel (tunable per plan)
// default outer panel of reduced-precision factorizations: on the 100^3
// 27-point operator (METIS) 256 brings the learned factor's backward error
// after one refinement step from 1.04e-12 to 8.2e-13 (under the 1e-12
// target: one solve less) for 2 ms more factor time
constexpr int kPanelOuterRed = 256;
constexpr int kPanelInner = 64;
constexpr int kTile = 128;
constexpr int kSyrkGroup = 8;  // Schur tiles per super-block side (one XCD's resident tiles)

}  // namespace smfgpu

using namespace smfgpu;

struct smfgpu_analysis_s {
    Analysis an;
};

struct smfgpu_plan_s {
    int32_t n = 0, nsup = 0;
    int32_t posdef = 1;
    int32_t device = 0;                            // HIP device current at plan creation
    hipStream_t stream = nullptr;
    // SPD: the factorization's launch sequence, captured once and replayed
    // (it reads the values from d_vals); [0]: every front in FP64, [1]:
    // with the reduced-precision Schur launches, [2]: with the learned
    // selector's launches too
    hipGraph_t fgraph[3] = {nullptr, nullptr, nullptr};
    hipGraphExec_t fexec[3] = {nullptr, nullptr, nullptr};
    int32_t graph_red_modes[3] = {-1, -1, -1};     // reduced Schur kernels each graph launches
    int32_t sched_red_min = INT32_MAX;             // SPD schedule: fronts with f >= this may run below FP64
    std::vector<FrontDev> fronts;
    std::vector<int32_t> level;                    // per supernode
    std::vector<std::vector<int32_t>> level_fronts;
    std::vector<int32_t> perm;
    int64_t nnz_perm = 0, n_orig = 0;
    int64_t factor_entries = 0, cb_entries = 0, rows_total = 0;
    double flops = 0.0;
    std::vector<double> front_flops;
    std::vector<Launch> launches;      // FP64 factorizations (outer panel panel_outer)
    std::vector<Launch> launches_red;  // reduced-precision ones (outer panel panel_outer_red)
    // device
    DBuf<FrontDev> d_fronts;
    DBuf<int32_t> d_rows, d_cmap, d_pcol_ptr, d_amap, d_perm, d_level_fronts, d_flag;
    DBuf<int64_t> d_cmofs, d_slot_ptr, d_slot_src;
    DBuf<double> d_av, d_L, d_CB, d_red, d_vals;
    DBuf<AsmTask> d_asm;
    DBuf<ChildCol> d_cc;
    DBuf<PotrfItem> d_potrf;
    DBuf<TrsmItem> d_trsm;
    DBuf<SyrkItem> d_syrk;
    DBuf<int8_t> d_prec;
    DBuf<double> d_junk;
    DBuf<uint16_t> d_xpl;  // large fronts' bf16 update operands (allocated when a BF16 / split-BF16 front needs them)
    DBuf<float> d_L32;  // allocated at the first factorization with FP32 Schur updates
    DBuf<uint16_t> d_L16h, d_L16l;  // packed bf16 planes, allocated (zeroed) at the first BF16 factorization
    int64_t l16_entries = 0;
    // solve schedule: items of every step of every level; step i of the
    // forward sweep is fwd_steps[i] = {first item, count}, level lv owns
    // steps [fwd_lv[lv], fwd_lv[lv + 1]) (backward likewise)
    DBuf<SolveItem> d_fwd_items, d_bwd_items;
    PinnedArr<SolveItem> h_fwd_items, h_bwd_items;  // LDL^T: staging of the solve items
    std::vector<std::pair<int64_t, int64_t>> fwd_steps, bwd_steps;
    std::vector<int32_t> fwd_lv, bwd_lv;
    // SPD sweeps as one persistent launch each: the items in the order the
    // sweep takes them (forward: as stored; backward: levels from the root
    // down), the step of each, the item count of each step, and the
    // counters (ticket, then one per step)
    DBuf<SolveItem> d_bwd_seq;
    DBuf<int32_t> d_fwd_step_of, d_bwd_step_of, d_fwd_step_cnt, d_bwd_step_cnt;
    DBuf<int32_t> d_sweep_ctr;
    DBuf<int32_t> d_sweep_err;  // set when a sweep's wait ran out (never expected)
    PinnedArr<int32_t> h_sweep_err;  // its host copy (pinned: read without an allocation)
    int32_t sweep_persistent = 0;  // per-step launches by default (measured faster)
    int32_t panel_outer = kPanelOuter;             // SPD: pivot columns per trailing Schur update
    int32_t panel_outer_red = kPanelOuterRed;      // the same for reduced-precision factorizations
    std::vector<int64_t> asm_first;                // SPD: first assembly task of each level
    int32_t cb_seg = 4096;                         // SPD: pivot columns per contribution-block update
    int32_t ldlt_big_min = 256;                    // LDL^T: fronts with at least this many rows take the large-front path
    DBuf<BigLd> d_big;                             // LDL^T: large-front states of the level
    int32_t ldlt_planes = 1;                       // LDL^T: bf16 operand planes for the large fronts' outer-block updates
    int32_t ldlt_outer_wide = 1024;                // LDL^T: outer block of the fronts updated on planes
    int32_t ldlt_group = 1;                        // LDL^T: speculative panel groups on the large-front path
    int32_t ldlt_update_wg = 1024;                 // LDL^T: workgroups of the large-front update launches
    DBuf<double> d_gpart;                          // LDL^T: per-workgroup column maxima of a group
    int32_t syrk_ahead = 2;                        // Schur update: chunks with global loads in flight (1 or 2)
    int32_t asm_chunk = 2048;                      // SPD assembly: largest LDS column chunk (rows)
    int32_t asm_wave_rows = 2048;                  // SPD assembly: levels of fronts up to this many rows take a wavefront per column
    int32_t solve_fuse_rows = 512;                 // rows below a last block its own workgroup updates
    double t_solve_phase[4] = {0, 0, 0, 0};        // fwd ms, bwd ms, fwd launches, bwd launches
    // refinement CSR (full symmetric, permuted numbering)
    DBuf<int64_t> d_fr_ptr, d_fr_slot;
    DBuf<int32_t> d_fr_col;
    // solve / refine work
    DBuf<double> d_y, d_b, d_r, d_rowstat, d_x, d_bb;
    // refinement: vectors of the iterations, per-column maxima, dot-product
    // scratch, values of a matrix other than the factored one, and
    // ||A||_inf of the factored values (valid until the next factorization)
    DBuf<double> d_rvec, d_stat, d_part, d_h, d_av_ref;
    double norm_a = 0.0;
    bool norm_a_valid = false;
    // symmetric scaling of the live factor: the scale (original numbering),
    // the unscaled permuted values (the refinement's residual matrix), and
    // the equilibration's two scale vectors
    DBuf<double> d_scale, d_av_raw, d_eqs;
    bool scaled = false;
    std::vector<int64_t> syrk_flops_of_launch, syrk_flops_of_launch_red;
    // learned selector: weights, per-column statistics of the assembled
    // pivot columns, per-front features, candidate fronts per level (SPD:
    // the K_SELECT launches' items)
    DBuf<float> d_selw, d_feat;
    DBuf<float4> d_colstat;
    DBuf<SelFront> d_selfr;
    std::vector<SelFront> selfr;
    bool sel_loaded = false;
    double sel_key[3] = {-1.0, -1.0, -1.0};         // tolerance, mask, scaling state baked into the graph [2]
    int32_t sel_fallbacks = 0;                      // refinements that fell back to an FP64 refactorization
    // LDL^T: per supernode precision code of the last factorization; the
    // training pairs' fronts and labels of a collecting factorization
    std::vector<int8_t> ldlt_prec;
    std::vector<int32_t> sel_cand;
    DBuf<unsigned int> d_lab;
    DBuf<double> d_snap;
    // timing
    bool timing = false;
    double t_syrk_ms = 0, t_asm_ms = 0, t_solve_ms = 0;
    double syrk_launches = 0, syrk_flops = 0, asm_bytes = 0, solve_bytes = 0;
    std::vector<hipEvent_t> ev, sev;
    int8_t last_prec_any = 0;
    smfgpu_factor_s *live = nullptr;
    std::set<smfgpu_factor_s *> issued;             // every factor handle of this plan not yet freed
    double max_abs = 0.0;
    // ---- LDL^T (posdef = 0) ----
    std::vector<int32_t> sn_parent, child_ptr, child_idx, front_row_ptr, front_row_idx;
    std::vector<int32_t> amap_h;                    // host copy of the symbolic-position map
    PinnedArr<FrontLd> fl;                          // host mirror of the fronts of the last factorization
    std::vector<int32_t> rows_h;                    // host mirror of the final row ids (rofs offsets)
    std::vector<int8_t> tags_h;
    DBuf<FrontLd> d_fl;
    DBuf<double> d_arena, d_work, d_dinv;
    DBuf<double> d_fac;                             // stored factor columns (f x ne per front)
    RangePool sq_pool;                              // live front squares in d_arena
    DBuf<int32_t> d_lrows, d_lcmap, d_flist;
    DBuf<int8_t> d_ltags, d_dkind;
    DBuf<unsigned char> d_stage;                    // per-level upload staging
    std::vector<unsigned char> h_stage;
    int64_t fac_used = 0, rows_used = 0, dinv_used = 0;
    std::vector<int64_t> lv_flist_ofs;              // per level: offset of its fronts in d_flist (solve)
    std::vector<int32_t> lv_fsym_max;               // per level: largest symbolic front
    std::vector<int64_t> lv_room;                   // prefix over levels of the device-level room estimates (4 per level)
    int32_t dev_fmax = kDevFrontMax;                // largest realized front of a device level (0: host layout only)
    int32_t dev_levels = 0;                         // levels the last factorization ran on the device
    int32_t dev_wave = 1;                           // device levels: one wavefront per front of <= 48 rows
    double t_ldlt_host[4] = {0, 0, 0, 0};           // s: device levels (run), their read-back, statistics, solve lists
    DBuf<int32_t> d_child_ptr, d_child_idx, d_parent;
    DBuf<unsigned long long> d_dctr;                // device-level counters
    unsigned long long h_dctr[kDcN] = {};
    bool ldlt_ready = false;                        // a factorization's fronts are valid
};

struct smfgpu_factor_s {
    smfgpu_plan_s *plan = nullptr;
    smfgpu_factor_info info{};
    bool valid = false;
    std::vector<int8_t> prec;                       // per supernode: device precision code of its Schur updates
    bool collected = false;                         // features kept for smfgpu_selector_pairs
    bool auto_prec = false;                         // precisions chosen by the learned selector
    int32_t mixed_min_front = 0;
};


namespace smfgpu {
// device precision code of a SMFGPU_PREC_* request (0 FP64, 1 FP32, 2 BF16,
// 3 split BF16), -1 when the request cannot be served
int prec_code(int32_t precision);
int ldlt_plan_build(smfgpu_plan_s *P, const smfgpu_analysis_desc *D);
int ldlt_factor(smfgpu_plan_s *P, const double *dv, int64_t nvalues, const smfgpu_factor_options &opt,
                smfgpu_factor_info &fi);
// job: SolveJob (types.hpp:250-256): 0 full, 1 forward (in: original
// order, out: permuted), 2 diagonal only, 3 backward, 4 diagonal then
// backward (in: permuted order; 3 and 4 out: original)
// ybuf (optional): caller-owned device block of ycap >= n doubles that
// replaces the plan's right-hand-side work block; the right-hand sides then
// go through in groups of at most ycap / n columns, nothing is allocated and
// nothing is timed
int ldlt_solve(smfgpu_plan_s *P, const double *b_dev, int64_t ldb, double *out_dev, int64_t ldo, int32_t nrhs,
               int32_t job, double *ybuf = nullptr, int64_t ycap = 0);
int ldlt_export(smfgpu_plan_s *P, int64_t *sizes, int64_t *steps, double *fv, int32_t *rows, int8_t *tags);
// solve on device buffers (b, out: n x nrhs, leading dimensions ldb / ldo;
// may alias), job as ldlt_solve, for either factor kind
int solve_device(smfgpu_plan_s *P, const double *b_dev, int64_t ldb, double *out_dev, int64_t ldo, int32_t nrhs,
                 int32_t job = 0, double *ybuf = nullptr, int64_t ycap = 0);
// after a synchronization: an error if a persistent sweep's dependency wait
// ran out since the last check (never expected; the flag is cleared)
int sweep_check(smfgpu_plan_s *P);
// device address of the no-allocation solve's work block on `device`,
// resolved once per device (noalloc.hip); the first call must be made with
// that device current
int scratch_init(int device);
// the selector launches' parameters for a factorization's options (w null
// unless the precision is AUTO)
SelPrm sel_prm(smfgpu_plan_s *P, const smfgpu_factor_options &opt);
}  // namespace smfgpu
