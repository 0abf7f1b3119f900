// Device-resident solver for the steady StaticBTESynthetic path.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "giftbte.h"
#include "ndchol.hpp"
#include "problem.hpp"

namespace gbte {

// Cells and nodes renumbered along a Morton curve (reorder.cpp).
Problem renumber(const Problem &p, std::vector<int> &cell_new, std::vector<int> &node_new);

// Per-mode constants handed to the kernels (device pointers and scalars).
struct Modes {
    const double *vg, *tau, *cap, *lr, *hr, *sx, *sy, *sz, *w;
    const double *lrc;  // lr / cap per band
    double kappa, cbulk;
    // Energy densities are held as deviations e - C_b eref from the
    // equilibrium at eref (the deck's Tref): the transport equations are
    // linear and every face sum of (A/V)(s.n) vanishes, so only the sources,
    // the thermalising walls and the moments see the shift.  off* are the
    // moments of C_b eref, added back to T, E, T_band, q and T_mac.
    double eref, offT, offE, offW, offQx, offQy, offQz, offMac;
};

// Owning device buffer.
template <typename T>
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
};

// Upwind record of one position of the wave-specialised sweep, 48 bytes:
// q >= 0 ring position, -1 none, -2-f far entry f of the sub-level; g the
// in-face (A/V)(s.n) and G the sum over out-faces.
struct PcRec {
    int cell, q[3];
    double g[3], G;
};

// Upwind record of one position of the level-synchronous chunked sweep
// (k_sweep_lv, tetrahedra), 48 bytes: the cell, its interior upwind cells
// (-1: none) with their in-face (A/V)(s.n), and G the sum over out-faces.
struct LvRec {
    int cell, c[3];
    double g[3], G;
};

// The wave-specialised sweep's schedule, as device pointers.
struct PcSched {
    const PcRec *rec;            // [dir][position]
    const unsigned short *size;  // per sub-level: (positions - 1) | far reads << 8, Lsub + 1 per direction
    const int *f0;               // per direction: offset of its first far entry in farcell
    const int *dsl;              // per direction offset into size
    const int *farcell;          // far upwind cells, per sub-level runs
};

// Compressed sparse rows of the synthetic (macro) and Fourier matrices,
// stored negated so that they are symmetric positive definite.
struct DevCSR {
    int n = 0;
    DBuf<int> ptr, col;
    DBuf<double> val, dinv;
};

// Nested-dissection Cholesky factor of the same matrix (ndchol.cpp) and the
// 64-row work items of each elimination-tree height (forward) and depth
// (backward).
struct DevND {
    int n = 0, max_front = 0, nodes = 0;
    DBuf<int> perm, col0, ncol, nrow, child_ptr, child_idx, map, map_off, rows, rows_off, u_off;
    DBuf<long> f_off;
    DBuf<double> frm, fcm;
    std::vector<int> fwd_ptr, bwd_ptr;  // level pointers into fwd_items / bwd_items
    std::vector<int> fwd_len, bwd_len;  // per level: the longest reduction of a row block
    DBuf<int2> fwd_items, bwd_items;    // (supernode, first row) per workgroup
    long u_total = 0;
};

class Solver {
   public:
    Solver(const Problem &p, const gbte_options &o);
    ~Solver();
    bool ok() const { return err_.empty(); }
    const std::string &error() const { return err_; }
    const std::string &tuning_log() const { return tune_log_; }

    int fourier(gbte_stats *st);
    int set_initial(const double *T, const double *qx, const double *qy, const double *qz);
    int solve(int max_iter, double tolT, double tolQ, gbte_stats *st);
    int get_fields(double *TL, double *TE, double *qx, double *qy, double *qz, double *Tb);
    int max_levels() const { return nlev; }
    int n_cells() const { return nc; }
    int n_bands() const { return nB; }
    int keff(int axis, double dT, double *out);
    // discrete adjoint of the converged state (adjoint.hip): dk_eff / d(tau
    // scale of each band) and dk_eff / d(temperature of each boundary face)
    int adjoint(int axis, double dT, int max_iter, double tol, double *dk_ds, double *dk_dtb, gbte_adjoint_stats *st);
    // record the synthetic-pass graphs now (the first solve does it
    // otherwise); batched solves do this before their threads start, as a
    // capture must not overlap another thread's allocations and copies
    int prepare();
    // one stream per solve and the sources in one pass (batched solves, which
    // keep the GPU busy by running side by side)
    void set_single_stream(bool on) { single_stream_ = on; }

    void set_error(const std::string &e) { err_ = e; }

   private:
    template <typename T>
    bool upload(DBuf<T> &b, const std::vector<T> &h);
    template <typename T>
    bool alloc(DBuf<T> &b, size_t n);
    bool check(hipError_t e, const char *what);
    bool build();
    // pre: the nested-dissection factor of the same rows, computed elsewhere
    // (with its error message in pre_err), or null to factor here
    bool build_matrix(const std::vector<std::vector<std::pair<int, double>>> &rows, DevCSR &m, DevND &f,
                      NDFactor *pre, const std::string *pre_err);
    // tetrahedral meshes (solver3d.hip): device tables, the solve loop and
    // the launches of the 3D kernels
    bool build3d(const std::vector<int> &be_slot);
    bool build_gv(double invLx);
    int solve3d(int max_iter, double tolT, double tolQ, gbte_stats *st);
    void launch_rhs3(bool macro, const double *bval, double *rhs);
    void launch_vsum3();
    void launch_grad3();
    double *sbuf();
    bool src_width_ok(int v) const;
    void launch_source3(const Modes &m);
    void launch_bound_out3(const Modes &m);
    void launch_moments3(const Modes &m);
    bool tune3d();
    void launch_sweep_pc(const PcSched &ps, const Modes &m, const double *tpos);
    void launch_sweep3(const Modes &m);
    void launch_sweep_lv(const Modes &m, int bc);
    Modes modes() const;
    // macro / Fourier linear solve of (-K) x = -rhs: the direct factor solve,
    // or PCG warm-started from x
    bool linsolve(const DevCSR &K, const DevND &f, double *d_rhs, double *d_x, int *iters);
    bool cg(const DevCSR &K, double *d_rhs, double *d_x, int *iters);
    bool nd_solve(const DevND &f, const double *d_rhs, double *d_x);
    bool reduce_host(const double *d_partials, int n, std::vector<double> &out, int stride);
    void launch_vertex_T(const double *T, double *TV);
    void launch_rhs(bool macro, const double *bval, double *rhs);
    bool inner_error(const double *a, const double *b, double *out);
    bool macro_launch(int q);
    bool macro_capture(int q);
    double macro_inner(int q) const;

    std::vector<int> cell_new_, node_new_;  // deck order -> device order
    Problem Pp;                             // renumbered problem (device order)
    const Problem &P;
    gbte_options opt{};
    std::string err_;
    hipStream_t s_ = nullptr;
    // before/after: source | sweep | bound_out | face kernels | boundary | macro loop
    hipEvent_t ev_[8] = {};
    // side stream: the next iteration's interior sources run there, after the
    // face kernel, beside the boundary update and the synthetic passes
    hipStream_t s2_ = nullptr;
    // per launch parity: gradients ready | interior sources start | boundary-cell sources end | boundary values
    // ready | interior sources end | boundary-cell sources start
    hipEvent_t ev2_[2][6] = {};
    bool single_stream_ = false;

    int nc = 0, nb = 0, ne = 0, nn = 0, nB = 0, nD = 0, nf = 0, dim = 2, cpw = 1, nlev = 0;
    int row_tpr = 256, row_rpb = 1;  // threads per cell row, rows per block
    int gnnz = 0;                    // padded LSQ stencil entries

    // mode tables
    DBuf<double> d_vg, d_tau, d_cap, d_lr, d_hr, d_sx, d_sy, d_sz, d_w, d_lrc;
    // cell / slot geometry
    DBuf<int> d_fsize, d_snb, d_sent;
    DBuf<double> d_vol, d_Q, d_snx, d_sny, d_snz, d_sarea, d_srx, d_sry, d_srz, d_sqx, d_sqy, d_sqz;
    // per-direction upwind schedule: positions in upwind order, level
    // pointers, in-face records (neighbour position, neighbour cell, g)
    DBuf<int> d_item_cell, d_posc, d_levptr, d_dlev, d_rec_q, d_rec_c;
    DBuf<int> d_src_int, d_src_bnd;  // cells without / with a boundary face
    int n_src_int = 0, n_src_bnd = 0;
    DBuf<double> d_tpos;             // [dir][position] T_old - Tref
    DBuf<double> d_rec_g, d_rec_G;
    int sweep_bc = 1, sweep_nchunk = 1, sweep_ring = 1;
    // wave-specialised sweep (k_sweep_pc): records, sub-level tables
    bool use_pc = false;
    int pc_bc = 1, pc_nchunk = 1, pc_ring = 1, nsub = 0, pc_far = 16;
    size_t pc_lds = 0;
    DBuf<PcRec> d_pc_rec;
    DBuf<int> d_pc_f0, d_pc_dsl, d_pc_far;
    DBuf<unsigned short> d_pc_size;
    // level-synchronous chunked sweep (k_sweep_lv, tetrahedra): records in
    // upwind order; lv_bc bands per workgroup when chosen (0: not used).
    // The sweep source S is then stored [dir][band chunk][position][lv_bc]
    // (s_sh = log2 lv_bc), else [dir][position][band] (s_sh = -1)
    DBuf<LvRec> d_lv_rec;
    int lv_bc = 0, lv_np = 2, s_sh = -1;
    int lv_v = 1, lv_tpb = 256;  // lanes per position (1: a lane takes the whole chunk), threads per workgroup
    int src_v = 0;  // FP32 3D source: bands per lane (2 or 4; 0: 2 where the layout takes it, else 1)
    // gradient stencil
    DBuf<int> d_gptr, d_gidx, d_gok;
    DBuf<double> d_gP;
    // IDW stencils (energy density: cells or boundary entries; temperature: cells)
    DBuf<int> d_vptr, d_vsrc, d_tptr, d_tsrc;
    DBuf<double> d_vwt, d_tdist;
    // face-gradient / non-Fourier slot data
    DBuf<int> d_fkind, d_fother, d_fv0, d_fv1, d_srei, d_slim;
    DBuf<double> d_srec;  // per-cell face records of the source kernel
    DBuf<double> d_fga, d_fn0x, d_fn0y, d_fmx, d_fmy;
    // boundary entries
    std::vector<int> h_ent_list;
    DBuf<int> d_ent_list, d_be_cell, d_be_type, d_be_conn, d_be_spec, d_be_slot, d_nint;
    DBuf<double> d_be_nx, d_be_ny, d_be_nz, d_be_tw, d_be_dT, d_be_tov;
    // macro / Fourier right-hand side slot data (two sets)
    DBuf<int> d_mbent[2], d_mnonorth[2], d_mv0[2], d_mv1[2];
    DBuf<double> d_mbcoef[2], d_mkA[2], d_ma2x[2], d_ma2y[2], d_mmx[2], d_mmy[2], d_mm2[2];
    DBuf<double> d_fourier_bval;
    DevCSR Kmacro, Kfour;
    // LDS tiles (tiles.cpp) and the host tables they are built from
    bool use_tiles = false;
    int n_tiles = 0, tile_rows = 0;
    size_t tile_lds = 0;
    DBuf<int> d_t_cell0, d_t_rptr, d_t_rsrc, d_t_nh, d_t_vptr, d_t_vs_ptr, d_t_vs_row, d_t_fin_i;
    DBuf<double> d_t_vs_w, d_t_cmf, d_t_kwf, d_t_fin_d;
    std::vector<int> h_gptr, h_gidx, h_vptr, h_vsrc, h_fkind, h_fother, h_fv0, h_fv1, h_gok;
    std::vector<double> h_vwt, h_gP, h_fga, h_fn0x, h_fn0y, h_fmxm, h_fmym;
    DevND NDmacro, NDfour;
    // 3D: source records, LSQ gradient tiles, moments, residual and
    // right-hand-side coefficients (solver3d.hip)
    DBuf<double> d_s3d, d_Y, d_YV, d_f3d, d_r3d[2];
    DBuf<unsigned char> d_oflag;  // 3D out-face flags per (cell, direction), bit per face
    DBuf<double> d_dirf;          // 3D moments: direction factors per direction
    DBuf<int> d_s3i, d_f3i, d_r3i[2];
    DBuf<char> d_g3g_ent;
    DBuf<int> d_g3g_ptr, d_g3g_idx;
    int g3_gs = 4, g3_wpb = 4;  // k_grad3g: 64-entry slices per wave, waves per workgroup (chosen by tune3d)
    int gv_gs = 4, gv_wpb = 4;  // k_vsum3: the same, chosen on its own
    // matrix-core gradient (k_grad3m): tiles of up to four cells, their rows,
    // cells and A' operands; used when the tuner finds it faster
    bool gm_ok = false, use_gm = false;
    int gm_ntiles = 0;
    DBuf<int> d_gm_rows, d_gm_cell;
    DBuf<char> d_gm_A;
    // vertex-sum gradient (k_vsum3 + k_grad3v): per-vertex entries, chosen
    // cells and sums, per-cell entries and records
    bool gv_ok = false, use_gv = false;
    int gv_nv = 0;
    DBuf<int> d_gv_vptr, d_gv_vidx, d_gv_vref, d_gv_ptr, d_gv_idx;
    DBuf<char> d_gv_vent, d_gv_ent, d_gv_rec, d_gv_sums;
    std::string tune_log_;  // the autotuner's candidates, timings and choices
    std::string phase3d_;   // build3d's host phase times
    DBuf<double> d_ndy, d_ndu, d_ndx;
    // state
    DBuf<double> d_e, d_grad, d_eb, d_eV, d_T, d_Told, d_T1, d_TV, d_E, d_qx, d_qy, d_qz, d_qxo, d_qyo, d_qzo,
        d_RM, d_Tb, d_tmac, d_rhs, d_S;
    // CG / reduction scratch
    DBuf<double> d_r, d_z, d_p, d_Ap, d_part, d_part2;
    double *h_pin = nullptr;  // pinned host scratch
    bool fp32_ = false;   // GBTE_FP32_SWEEP: per-system fields stored in float
    double eref_ = 300.0;  // energy densities are held as e - C_b eref_
    double offT_ = 0, offE_ = 0, offW_ = 0, offQ_[3] = {0, 0, 0}, offMac_ = 0;  // moments of C_b eref_
    double *h_part = nullptr;  // pinned copies of the synthetic passes' change partials (2 slots)
    double *h_err = nullptr;   // pinned copy of the outer-error partials
    DBuf<double> d_mpart;      // the passes' change partials (2 slots)
    hipGraph_t macro_graph_[2] = {nullptr, nullptr};
    hipGraphExec_t macro_exec_[2] = {nullptr, nullptr};
    int red_blocks = 0;
    bool have_state = false;
    double errInner_ = 0, errTold_ = 0, errQold_ = 0;
    int errInc_ = 0;
    int fourier_passes_ = 0;
    int last_passes_ = 0;  // synthetic passes of the last outer iteration
};

template <typename T>
bool Solver::alloc(DBuf<T> &b, size_t n) {
    b.release();
    b.n = n;
    if (n == 0) return true;
    return check(hipMalloc((void **)&b.p, n * sizeof(T)), "hipMalloc");
}

template <typename T>
bool Solver::upload(DBuf<T> &b, const std::vector<T> &h) {
    if (!alloc(b, h.size())) return false;
    if (h.empty()) return true;
    return check(hipMemcpy(b.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
}

}  // namespace gbte
