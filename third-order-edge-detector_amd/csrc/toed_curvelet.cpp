// Curvelet formation over a TOED edgel list -- host C++, part of libtoed_hip.so.
//
// The consumer the reference runs after its FP64 detector
// (main_gpu_cpu.cpp:211-281, main_cpu_DP.cpp:208-281): the N x 4 list
// [x, y, orientation, strength] is grouped into curvelets (anchored groups of
// up to max_size_to_group edgels that one curve bundle fits), the curvelets
// are linked into an edgel link graph, unambiguous chains are extracted as
// curve fragments and poly-arcs are fitted to the fragments
// (curvelet/form_curvelet_process.hpp:125-163 execute()).  Output types:
//   0  curvelet map   (curveletmap.hpp get_output_array, curvelet.hpp set_output)
//   1  curve fragments (curve_fragment_graph.hpp get_output_array)
//   2  poly-arcs      (polyarcmap.hpp get_output_array)
// Outputs are the reference's column-major arrays (Array.hpp val(r,c) =
// data[c*h + r]); cells the reference leaves unwritten are 0 here.
//
// Results are bit-identical to the reference's compiled code: every
// floating-point expression is evaluated in the reference's order, with the
// same libm calls on the same arguments, and hypotheses are ordered by the
// same std::sort over the same sequence.  What differs is the organisation:
//  * the edgel map is a CSR grid of W x H cells instead of a std::map of
//    vectors, filled in id order (the reference's insertion order);
//  * every sine and cosine that depends on one edgel only (theta +- dt,
//    theta + Dt(j), theta) is evaluated once per edgel instead of once per
//    bundle cell; only the sines of the pairwise angle differences remain per
//    pair and column (2 x 7 calls for the default 9 x 7 bundle grid instead
//    of 6 x 63);
//  * a neighbour is classified forward / backward once (the two grouping
//    passes share the neighbourhood scan and the direction test);
//  * anchors are independent in the grouping stage (a duplicate test only
//    looks at the anchor's own curvelets), so that stage runs on a thread
//    pool over blocks of anchors, merged in anchor order;
//  * sine / cosine pairs of one argument: the reference's compiled code
//    (g++ -O3 on x86-64, the flags of its makefile) evaluates some of them
//    with glibc's sincos and others with separate sin and cos calls, and the
//    two can differ in the last place.  Each pair here is evaluated the way
//    that code evaluates it (sc_pair / sep_pair below), and this file is
//    compiled with -fno-builtin-sin -fno-builtin-cos so that the compiler
//    does not merge the separate calls itself;
//  * poly-arc fitting runs per fragment in parallel; the reference re-applies
//    the end-point refit to every arc found so far after each fragment
//    (form_curvelet_process.hpp fit_polyarc_to_chain ->
//    polyarcmap.hpp compute_best_fit_according_to_curvelets), which is
//    quadratic.  That refit depends on the arc's state only through whether
//    |k| < k_th on entry, and after two applications that test no longer
//    changes, so applying it min(F - f, 2) times to the arcs of fragment f
//    (of F) is the same computation.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

#include "toed_capi.h"

namespace {

constexpr double kPi = 3.14159265358979323846;      // curvelet_utils.hpp M_PI
constexpr double kKth = 1e-4;                        // curvelet_utils.hpp k_th
constexpr double kStepX = 0.1;                       // CC_curve_model_3d.hpp _dx_ (pixels)
constexpr double kStepT = 0.08;                      // CC_curve_model_3d.hpp _dt_ (radians)
constexpr unsigned kMinLinkGroup = 4;                // form_curvelet_process.hpp execute(): min_size_to_link

struct Pt {
    double x, y;
};

inline double sq_dist(const Pt &a, const Pt &b) {    // curvelet_utils.hpp sq_distance
    return (a.x - b.x) * (a.x - b.x) + (a.y - b.y) * (a.y - b.y);
}

inline double sq_norm(const Pt &a) { return a.x * a.x + a.y * a.y; }

// sin and cos of one argument through glibc's sincos, or through separate
// calls (see the file comment: the choice follows the reference's build)
inline void sc_pair(double x, double &s, double &c) { ::sincos(x, &s, &c); }
inline void sep_pair(double x, double &s, double &c) {
    s = ::sin(x);
    c = ::cos(x);
}

// curvelet_utils.hpp To2Pi: wrap into [0, 2 pi), anything else -> 0
inline double wrap_2pi(double a) {
    double r;
    if (a >= 2 * kPi)
        r = std::fmod(a, 2 * kPi);
    else if (a < 0)
        r = 2 * kPi + std::fmod(a, 2 * kPi);
    else
        r = a;
    if (!(r >= 0 && r < 2 * kPi)) r = 0;
    return r;
}

struct Params {
    double rad, gap, dx, dt, token_len, max_k;
    unsigned group;          // max_size_to_group
    bool centered, bidirectional;
};

// One curve bundle: lower and upper curvature surfaces over the (Dx, Dt)
// grid, stored column-major [j * ni + i] like the reference's Array; only
// best_fit depends on an iteration order (i outer, j inner, as there).
struct BundleRef {
    double *kmin, *kmax;
};

struct Curvelet {
    int anchor;
    int n;                   // chain length
    int ids;                 // offset of the chain in the id pool
    double info[10];
};

struct Arc {
    int n, ids;
    double info[12];
    bool poly;
};

// One Dt column of a bundle (CC_curve_model_3d.hpp compute_curve_bundle's
// cell body, both branches): cells i = 0..ni-1 start from [-max_k, max_k]
// and are bounded by the two position bounds and the two orientation bounds
// in the reference's order (the order decides which of two equal bounds of
// opposite zero sign is kept).  cxm, sym / cxp, syp are DX cos(theta_b -+ dt)
// and DY sin(theta_b -+ dt); each denominator is summed in the reference's
// order, (-s dx2 + cx) + sy.  Plain IEEE arithmetic, so the SIMD clones give
// the same bits as the scalar code.
__attribute__((target_clones("avx512f", "avx2", "default")))
void bound_column(bool behind, const double *__restrict DXp, unsigned ni, double *__restrict lo_j,
                  double *__restrict hi_j, double tt, double DX2, double DY2, double dx_sq, double dx, double mk,
                  double s_m, double s_p, double cxm, double sym, double cxp, double syp) {
    if (behind) {
        for (unsigned i = 0; i < ni; ++i) {
            const double dx2 = DXp[i];
            const double ttmdx2 = tt - dx2;
            double lo = -mk, hi = mk;
            const double denom = (-2.0 * dx2 * tt + DX2 + DY2 - dx_sq + dx2 * dx2);
            const double k_dx_min = 2.0 * (ttmdx2 - dx) / denom;
            lo = lo < k_dx_min ? k_dx_min : lo;
            const double k_dx_max = 2.0 * (ttmdx2 + dx) / denom;
            hi = hi > k_dx_max ? k_dx_max : hi;
            const double k_dt_min = s_m / (-s_m * dx2 + cxm + sym);
            lo = lo < k_dt_min ? k_dt_min : lo;
            const double k_dt_max = s_p / (-s_p * dx2 + cxp + syp);
            hi = hi > k_dt_max ? k_dt_max : hi;
            lo_j[i] = lo;
            hi_j[i] = hi;
        }
    } else {
        for (unsigned i = 0; i < ni; ++i) {
            const double dx2 = DXp[i];
            const double ttmdx2 = tt - dx2;
            double lo = -mk, hi = mk;
            const double denom = -(2.0 * dx2 * tt - DX2 - DY2 + dx_sq - dx2 * dx2) / 2.0;
            const double k_dx_min = (ttmdx2 - dx) / denom;
            lo = lo < k_dx_min ? k_dx_min : lo;
            const double k_dx_max = (ttmdx2 + dx) / denom;
            hi = hi > k_dx_max ? k_dx_max : hi;
            const double k_dt_max = s_m / (-s_m * dx2 + cxm + sym);
            hi = hi > k_dt_max ? k_dt_max : hi;
            const double k_dt_min = s_p / (-s_p * dx2 + cxp + syp);
            lo = lo < k_dt_min ? k_dt_min : lo;
            lo_j[i] = lo;
            hi_j[i] = hi;
        }
    }
}

// CC_curve_model_3d.hpp intersect: Kmin = Array::array_max(Kmin1, Kmin2)
// (std::max(A1, A2)), Kmax = Array::array_min(Kmax1, Kmax2) (std::min(A1, A2)).
__attribute__((target_clones("avx512f", "avx2", "default")))
void intersect_cells(const double *__restrict lo1, const double *__restrict lo2, const double *__restrict hi1,
                     const double *__restrict hi2, double *__restrict lo, double *__restrict hi, unsigned nc) {
    for (unsigned c = 0; c < nc; ++c) {
        lo[c] = lo1[c] < lo2[c] ? lo2[c] : lo1[c];
        hi[c] = hi2[c] < hi1[c] ? hi2[c] : hi1[c];
    }
}

class Former {
public:
    Former(const double *cols, unsigned n, unsigned nrows, unsigned ncols, const Params &p)
        : P(p), N(n), H(nrows), W(ncols) {
        pt.resize(N);
        th.resize(N);
        for (unsigned i = 0; i < N; ++i) {          // edgemap.hpp:46-50: columns x, y, orientation
            pt[i] = Pt{cols[i], cols[(size_t)N + i]};
            th[i] = cols[2 * (size_t)N + i];
        }
        // CC_curve_model_3d.hpp compute_curve_bundle: grid of
        // (2 floor(dx/_dx_ + 1/2) + 1) x (2 floor(dt/_dt_ + 1/2) + 1) cells
        ni = (unsigned)(2 * std::floor(P.dx / kStepX + 0.5) + 1);
        nj = (unsigned)(2 * std::floor(P.dt / kStepT + 0.5) + 1);
        nc = ni * nj;
        DX.resize(ni);
        DT.resize(nj);
        for (unsigned i = 0; i < ni; ++i) DX[i] = kStepX * ((double)i - (double(ni) - 1.0) / 2.0);
        for (unsigned j = 0; j < nj; ++j) DT[j] = kStepT * ((double)j - (double(nj) - 1.0) / 2.0);
        nrad = (unsigned)std::ceil(P.rad);
        rad_sqr = (unsigned)(P.rad * P.rad);         // the reference keeps it in an unsigned
    }

    void execute(int nthreads);

    void output_size(unsigned type, unsigned &h, unsigned &w) const {
        if (type == 0) {
            h = (unsigned)cvlets.size();
            w = P.group + 1;
        } else if (type == 1) {
            h = (unsigned)frag_off.size() - 1;
            w = frag_maxlen;
        } else {
            h = (unsigned)arcs.size();
            w = arc_maxlen;
        }
    }

    void output(unsigned type, int *chain, double *info) const;

private:
    Params P;
    unsigned N, H, W;
    std::vector<Pt> pt;
    std::vector<double> th;
    unsigned ni = 0, nj = 0, nc = 0, nrad = 0, rad_sqr = 0;
    std::vector<double> DX, DT;
    // per-edgel trigonometry: cos, sin of theta; sin/cos of theta + dt and
    // theta - dt both through sincos ([0]) and through separate calls ([1]);
    // sin/cos of theta + Dt(j)
    std::vector<double> e_cos, e_sin, e_sj, e_cj;
    std::vector<double> e_sp[2], e_cp[2], e_sm[2], e_cm[2];
    // CSR grid of cells (x = round(x), y = round(y)) inside the image
    std::vector<int> cell_start, cell_ids;

    // stage outputs
    std::vector<Curvelet> cvlets;
    std::vector<int> cv_ids;
    std::vector<int> frag_off, frag_ids;
    unsigned frag_maxlen = 0;
    std::vector<Arc> arcs;
    std::vector<int> arc_ids;
    unsigned arc_maxlen = 0;

    int nthreads_ = 1;

    template <class F>
    void parallel_blocks(size_t n, size_t block, F &&fn);

    void precompute();
    bool bundle(int a, int b, BundleRef out) const;
    bool valid(const BundleRef &b) const {
        for (unsigned c = 0; c < nc; ++c)
            if (b.kmax[c] > b.kmin[c]) return true;
        return false;
    }
    bool intersect(const BundleRef &m1, const BundleRef &m2, BundleRef out) const {
        intersect_cells(m1.kmin, m2.kmin, m1.kmax, m2.kmax, out.kmin, out.kmax, nc);
        return valid(out);
    }
    int dir_sign(int a, int b, double d2, bool centered) const;
    void best_fit(int anchor, const BundleRef &b, double *info) const;
    void properties(const int *chain, int n, int anchor, double &length, double &quality) const;

    void group_anchor(int a, std::vector<Curvelet> &out, std::vector<int> &ids, std::vector<double> &scratch,
                      std::vector<int> &work) const;
    void stage_groups();
    void stage_links_and_fragments();
    void stage_polyarcs();
};

template <class F>
void Former::parallel_blocks(size_t n, size_t block, F &&fn) {
    const size_t nb = (n + block - 1) / block;
    const int nt = (int)std::min<size_t>((size_t)std::max(1, nthreads_), std::max<size_t>(nb, 1));
    std::atomic<size_t> next{0};
    auto worker = [&](int t) {
        for (;;) {
            const size_t b = next.fetch_add(1);
            if (b >= nb) break;
            fn(t, b, b * block, std::min(n, (b + 1) * block));
        }
    };
    if (nt <= 1) {
        worker(0);
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve(nt - 1);
    for (int t = 1; t < nt; ++t) pool.emplace_back(worker, t);
    worker(0);
    for (auto &th_ : pool) th_.join();
}

void Former::precompute() {
    e_cos.resize(N);
    e_sin.resize(N);
    for (int v = 0; v < 2; ++v) {
        e_sp[v].resize(N);
        e_cp[v].resize(N);
        e_sm[v].resize(N);
        e_cm[v].resize(N);
    }
    e_sj.resize((size_t)N * nj);
    e_cj.resize((size_t)N * nj);
    const double dt = P.dt;
    parallel_blocks(N, 4096, [&](int, size_t, size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i) {
            const double t = th[i];
            sc_pair(t, e_sin[i], e_cos[i]);
            const double tp = t + dt, tm = t - dt;
            sc_pair(tp, e_sp[0][i], e_cp[0][i]);
            sc_pair(tm, e_sm[0][i], e_cm[0][i]);
            sep_pair(tp, e_sp[1][i], e_cp[1][i]);
            sep_pair(tm, e_sm[1][i], e_cm[1][i]);
            for (unsigned j = 0; j < nj; ++j) sc_pair(t + DT[j], e_sj[i * nj + j], e_cj[i * nj + j]);
        }
    });
    // edgemap.hpp insert(): key (round(x), round(y)); cells outside the image
    // are never visited by the neighbourhood scan, so they are not stored
    cell_start.assign((size_t)W * H + 1, 0);
    std::vector<int> cell_of(N, -1);
    for (unsigned i = 0; i < N; ++i) {
        const double rx = std::round(pt[i].x), ry = std::round(pt[i].y);
        if (rx >= 0 && ry >= 0 && rx < (double)W && ry < (double)H) {
            const int c = (int)ry * (int)W + (int)rx;
            cell_of[i] = c;
            ++cell_start[c + 1];
        }
    }
    for (size_t c = 0; c < (size_t)W * H; ++c) cell_start[c + 1] += cell_start[c];
    cell_ids.resize(cell_start[(size_t)W * H]);
    std::vector<int> fill(cell_start.begin(), cell_start.end() - 1);
    for (unsigned i = 0; i < N; ++i)
        if (cell_of[i] >= 0) cell_ids[fill[cell_of[i]]++] = (int)i;     // id order within a cell
}

// CC_curve_model_3d.hpp compute_curve_bundle(ref_e = a, ne = b, dx, dt, max_k).
// Returns whether the bundle is valid (some cell with kmax > kmin).
bool Former::bundle(int a, int b, BundleRef out) const {
    const double mk = P.max_k;
    const double dx = P.dx;
    const double DXv = pt[a].x - pt[b].x;
    const double DYv = pt[a].y - pt[b].y;
    const double DX2 = DXv * DXv;
    const double DY2 = DYv * DYv;
    if (DX2 + DY2 < 4 * dx * dx) {
        for (unsigned c = 0; c < nc; ++c) {
            out.kmin[c] = -mk;
            out.kmax[c] = mk;
        }
        return nc > 0 && mk > -mk;
    }
    const double dx_sq = dx * dx;
    const bool behind = (DXv * e_cos[a] + DYv * e_sin[a]) < 0;
    // theta_b +- dt: sincos on the `behind` side, separate calls on the other
    const int v = behind ? 0 : 1;
    const double sin_T0pdt_times_DY = DYv * e_sp[v][b];
    const double cos_T0pdt_times_DX = DXv * e_cp[v][b];
    const double sin_T0mdt_times_DY = DYv * e_sm[v][b];
    const double cos_T0mdt_times_DX = DXv * e_cm[v][b];
    const double T0_p_dt = th[b] + P.dt;
    const double T0_m_dt = th[b] - P.dt;
    const double T2 = th[a];
    const double *sj = &e_sj[(size_t)a * nj], *cj = &e_cj[(size_t)a * nj];
    for (unsigned j = 0; j < nj; ++j) {
        const double T2t = T2 + DT[j];
        const double tt = sj[j] * DXv - cj[j] * DYv;
        const double s_m = ::sin(T2t - T0_m_dt);
        const double s_p = ::sin(T2t - T0_p_dt);
        bound_column(behind, DX.data(), ni, out.kmin + (size_t)j * ni, out.kmax + (size_t)j * ni, tt, DX2, DY2,
                     dx_sq, dx, mk, s_m, s_p, cos_T0mdt_times_DX, sin_T0mdt_times_DY, cos_T0pdt_times_DX,
                     sin_T0pdt_times_DY);
    }
    return valid(out);
}

// The sign of form_curvelet_process.hpp form_a_hypothesis's
// dot(vPointPoint(a, b), theta_a) = cos(dir) cos(theta_a) + sin(dir) sin(theta_a),
// dir = To2Pi(atan2(...)): +1, -1 or 0.  The reference's value is within
// 1e-14 of cos(dir - theta_a) = (DX cos theta_a + DY sin theta_a) / r, so
// where that projection exceeds 1e-12 r its sign is the reference's sign and
// the atan2 and the sine / cosine pair are skipped; otherwise the reference's
// expression is evaluated (sincos, or separate calls where `centered` is set:
// the reference's compiled form_a_hypothesis does so for the anchor-centred
// grouping; its poly-arc fit always passes centered = false).
int Former::dir_sign(int a, int b, double d2, bool centered) const {
    const double vx = pt[b].x - pt[a].x, vy = pt[b].y - pt[a].y;
    const double proj = vx * e_cos[a] + vy * e_sin[a];
    if (proj * proj > 1e-24 * d2) return proj > 0 ? 1 : -1;
    const double ref_dir = wrap_2pi(std::atan2(vy, vx));
    double sr, cr;
    if (centered)
        sep_pair(ref_dir, sr, cr);
    else
        sc_pair(ref_dir, sr, cr);
    const double d = cr * e_cos[a] + sr * e_sin[a];
    return d > 0 ? 1 : d < 0 ? -1 : 0;
}

// form_curvelet_process.hpp form_a_hypothesis: whether a neighbour on side
// `sgn` of the anchor forms a hypothesis in this pass, and on which side of
// the anchor it joins the chain (ref_first: after it).
static bool hyp_rule(int sgn, bool forward, bool centered, bool leading, bool &ref_first) {
    if (centered) {
        ref_first = (sgn > 0) == forward;
        return true;
    }
    if (sgn > 0 && forward && leading) {
        ref_first = true;
        return true;
    }
    if (sgn < 0) {
        if (forward && !leading) {
            ref_first = false;
            return true;
        }
        if (!forward) {
            ref_first = true;
            return true;
        }
    }
    return false;
}

// CC_curve_model_3d.hpp compute_best_fit + set_best_fit + set_output: writes
// info[1..7] = Kmax, Kmin at the best cell, ref orientation, position,
// orientation and curvature of the best fit.
void Former::best_fit(int anchor, const BundleRef &b, double *info) const {
    unsigned mini = ni, minj = nj;
    double mind = 100;
    for (unsigned i = 0; i < ni; i++)
        for (unsigned j = 0; j < nj; j++) {
            const unsigned c = j * ni + i;
            if (b.kmax[c] > b.kmin[c]) {
                const double d = DX[i] * DX[i] + DT[j] * DT[j];
                if (d < mind) {
                    mind = d;
                    mini = i;
                    minj = j;
                }
            }
        }
    const unsigned c = minj * ni + mini;
    const double dxb = DX[mini], dtb = DT[minj];
    const double theta = wrap_2pi(th[anchor] + dtb);
    info[1] = b.kmax[c];
    info[2] = b.kmin[c];
    info[3] = th[anchor];
    double st, ct;
    sc_pair(theta, st, ct);
    info[4] = pt[anchor].x - dxb * st;
    info[5] = pt[anchor].y + dxb * ct;
    info[6] = theta;
    info[7] = (b.kmax[c] + b.kmin[c]) / 2;
}

// curvelet.hpp compute_properties: length = (segments up to the anchor) +
// (segments after it), quality = 2 / (R / L + L / token_len / n)
void Former::properties(const int *chain, int n, int anchor, double &length, double &quality) const {
    double Lm = 0, Lp = 0;
    bool before = true;
    for (int i = 0; i + 1 < n; i++) {
        const double s = std::sqrt(sq_dist(pt[chain[i]], pt[chain[i + 1]]));
        if (before)
            Lm += s;
        else
            Lp += s;
        if (chain[i + 1] == anchor) before = false;
    }
    length = Lm + Lp;
    quality = 2 / (1.0 * P.rad / length + 1.0 * length / P.token_len / (double)n);
}

struct Hyp {
    double d;                // squared distance to the anchor
    int nb;                  // neighbour id
    int slot;                // bundle slot in the scratch arena
    bool ref_first;
};

// form_curvelet_process.hpp build_curvelets_greedy (one anchor, every pass it
// runs for the configured style) and build_curvelets_greedy_for_edge
void Former::group_anchor(int a, std::vector<Curvelet> &out, std::vector<int> &ids, std::vector<double> &scratch,
                          std::vector<int> &work) const {
    // passes, in the reference's call order: (forward, leading)
    bool pf[2], pl[2];
    int npass;
    if (P.centered) {
        pf[0] = true;  pl[0] = false;
        pf[1] = false; pl[1] = false;
        npass = P.bidirectional ? 2 : 1;
    } else if (P.bidirectional) {
        pf[0] = true;  pl[0] = true;
        pf[1] = false; pl[1] = true;
        npass = 2;
    } else {
        pf[0] = true; pl[0] = true;
        pf[1] = true; pl[1] = false;
        npass = 2;
    }
    const size_t first_out = out.size();
    const long ii = (long)std::round(pt[a].x), jj = (long)std::round(pt[a].y);
    const long nr = (long)nrad;
    const size_t G = std::max<unsigned>(P.group, 2);
    // neighbourhood scan, shared by the passes: x outer, y inner, id order
    // within a cell (the reference's scan order); each neighbour's side of
    // the anchor and, on first use, its bundle (the same in every pass)
    struct Nb {
        double d;
        int b, sgn, slot;       // slot: -2 not computed, -1 invalid bundle
    };
    std::vector<Nb> nbs;
    nbs.reserve(64);
    for (long xx = ii - nr; xx <= ii + nr; xx++) {
        if (xx < 0 || xx >= (long)W) continue;
        for (long yy = jj - nr; yy <= jj + nr; yy++) {
            if (yy < 0 || yy >= (long)H) continue;
            const size_t cell = (size_t)yy * W + (size_t)xx;
            for (int k = cell_start[cell]; k < cell_start[cell + 1]; ++k) {
                const int b = cell_ids[k];
                if (b == a) continue;
                const double d = sq_dist(pt[a], pt[b]);
                if (d > (double)rad_sqr) continue;
                nbs.push_back(Nb{d, b, dir_sign(a, b, d, P.centered), -2});
            }
        }
    }
    int slots = 0;
    for (int pass = 0; pass < npass; ++pass) {
        const bool forward = pf[pass], leading = pl[pass];
        std::vector<Hyp> hyps;
        hyps.reserve(nbs.size());
        for (Nb &nb : nbs) {
            bool ref_first = false;
            if (!hyp_rule(nb.sgn, forward, P.centered, leading, ref_first)) continue;
            if (nb.slot == -2) {
                if (scratch.size() < (size_t)(slots + 1) * 2 * nc) scratch.resize((size_t)(slots + 1) * 4 * nc);
                BundleRef br{&scratch[(size_t)slots * 2 * nc], &scratch[(size_t)slots * 2 * nc + nc]};
                nb.slot = bundle(a, nb.b, br) ? slots++ : -1;
            }
            if (nb.slot >= 0) hyps.push_back(Hyp{nb.d, nb.b, nb.slot, ref_first});
        }
        std::sort(hyps.begin(), hyps.end(), [](const Hyp &h1, const Hyp &h2) { return h1.d < h2.d; });
        const size_t nh = hyps.size();
        // two working bundles after the hypothesis slots
        if (scratch.size() < (size_t)(slots + 2) * 2 * nc) scratch.resize((size_t)(slots + 2) * 2 * nc);
        work.resize(2 * G + 4);
        for (size_t h1 = 0; h1 < nh; h1++) {
            // the chain grows either way from the middle of `work`
            int head = (int)G + 1, tail = head;
            work[tail++] = a;
            BundleRef cur{&scratch[(size_t)hyps[h1].slot * 2 * nc], &scratch[(size_t)hyps[h1].slot * 2 * nc + nc]};
            int flip = 0;
            for (size_t h2 = 0; h2 < nh; h2++) {
                if (h2 == h1) {
                    if (hyps[h1].ref_first)
                        work[tail++] = hyps[h1].nb;
                    else
                        work[--head] = hyps[h1].nb;
                    if ((size_t)(tail - head) >= P.group)
                        break;
                    else
                        continue;
                }
                double *nb = &scratch[(size_t)(slots + flip) * 2 * nc];
                BundleRef cand{nb, nb + nc};
                const BundleRef other{&scratch[(size_t)hyps[h2].slot * 2 * nc],
                                      &scratch[(size_t)hyps[h2].slot * 2 * nc + nc]};
                if (intersect(cur, other, cand)) {
                    cur = cand;
                    flip ^= 1;
                    if (hyps[h2].ref_first)
                        work[tail++] = hyps[h2].nb;
                    else
                        work[--head] = hyps[h2].nb;
                }
                if ((size_t)(tail - head) >= P.group) break;
            }
            const int n = tail - head;
            if (n <= 2) continue;
            // curveletmap.hpp does_curvelet_exist: same chain already anchored here
            bool exists = false;
            for (size_t q = first_out; q < out.size() && !exists; ++q) {
                if (out[q].n != n) continue;
                exists = std::equal(work.begin() + head, work.begin() + tail, ids.begin() + out[q].ids);
            }
            if (exists) continue;
            Curvelet cv;
            cv.anchor = a;
            cv.n = n;
            cv.ids = (int)ids.size();
            ids.insert(ids.end(), work.begin() + head, work.begin() + tail);
            cv.info[0] = forward ? 1.0 : 0.0;
            best_fit(a, cur, cv.info);
            properties(&work[head], n, a, cv.info[8], cv.info[9]);
            out.push_back(cv);
        }
    }
}

void Former::stage_groups() {
    const size_t block = 256;
    const size_t nb = (N + block - 1) / block;
    std::vector<std::vector<Curvelet>> part(nb);
    std::vector<std::vector<int>> part_ids(nb);
    std::vector<std::vector<double>> scratch(std::max(1, nthreads_));
    std::vector<std::vector<int>> work(std::max(1, nthreads_));
    parallel_blocks(N, block, [&](int t, size_t b, size_t lo, size_t hi) {
        for (size_t a = lo; a < hi; ++a) group_anchor((int)a, part[b], part_ids[b], scratch[t], work[t]);
    });
    size_t total = 0, total_ids = 0;
    for (size_t b = 0; b < nb; ++b) {
        total += part[b].size();
        total_ids += part_ids[b].size();
    }
    cvlets.clear();
    cvlets.reserve(total);
    cv_ids.clear();
    cv_ids.reserve(total_ids);
    for (size_t b = 0; b < nb; ++b) {
        const int base = (int)cv_ids.size();
        for (Curvelet c : part[b]) {
            c.ids += base;
            cvlets.push_back(c);
        }
        cv_ids.insert(cv_ids.end(), part_ids[b].begin(), part_ids[b].end());
    }
}

// form_curvelet_process.hpp construct_the_link_graph + edgel_link_graph.hpp
// link_edgels, then extract_one_chains_from_the_link_graph.  Only the first
// child / parent link of an edgel and the link counts are read downstream.
void Former::stage_links_and_fragments() {
    struct Link {
        int pe, ce, next_child;
    };
    std::vector<Link> links;
    std::vector<int> c_head(N, -1), c_tail(N, -1), c_cnt(N, 0), p_first(N, -1), p_cnt(N, 0);
    auto find_child = [&](int e1, int e2) {
        for (int l = c_head[e1]; l >= 0; l = links[l].next_child)
            if (links[l].ce == e2) return true;
        return false;
    };
    for (const Curvelet &cv : cvlets) {              // anchors in order, their curvelets in order
        if ((unsigned)cv.n < kMinLinkGroup) continue;
        const int *ch = &cv_ids[cv.ids];
        for (int k = 0; k + 1 < cv.n; k++) {
            const int e1 = ch[k], e2 = ch[k + 1];
            if (find_child(e1, e2) || find_child(e2, e1)) continue;   // vote only
            const int l = (int)links.size();
            links.push_back(Link{e1, e2, -1});
            if (c_tail[e1] >= 0)
                links[c_tail[e1]].next_child = l;
            else
                c_head[e1] = l;
            c_tail[e1] = l;
            ++c_cnt[e1];
            if (p_first[e2] < 0) p_first[e2] = l;
            ++p_cnt[e2];
        }
    }
    auto child = [&](int e) { return links[c_head[e]].ce; };
    auto parent = [&](int e) { return links[p_first[e]].pe; };
    std::vector<char> linked(N, 0);
    auto legal_first = [&](int e) {
        return !linked[e] && (p_cnt[e] == 1 && c_cnt[e] == 1) && parent(e) != child(e);
    };
    auto legal = [&](int e) { return !linked[e] && p_cnt[e] <= 1 && c_cnt[e] <= 1; };
    const double thres = P.gap * P.gap;
    frag_off.assign(1, 0);
    frag_ids.clear();
    frag_maxlen = 0;
    std::vector<int> buf(2 * (size_t)N + 4);
    for (unsigned i = 0; i < N; i++) {
        const int first = (int)i;
        if (linked[first]) continue;
        if (!legal_first(first)) continue;
        int head = (int)N + 1, tail = head;
        buf[tail++] = first;
        linked[first] = 1;
        int eA = child(first), eB;
        if (sq_dist(pt[first], pt[eA]) < thres) buf[tail++] = eA;
        while (legal(eA)) {
            linked[eA] = 1;
            if (c_cnt[eA] == 0) break;
            eB = eA;
            eA = child(eA);
            if (sq_dist(pt[eB], pt[eA]) < thres)
                buf[tail++] = eA;
            else
                break;
        }
        if (eA != first) {
            eA = parent(first);
            if (sq_dist(pt[first], pt[eA]) < thres) buf[--head] = eA;
        }
        while (legal(eA)) {
            linked[eA] = 1;
            if (p_cnt[eA] == 0) break;
            eB = eA;
            eA = parent(eA);
            if (sq_dist(pt[eB], pt[eA]) < thres)
                buf[--head] = eA;
            else
                break;
        }
        const int n = tail - head;
        if (n > 2) {
            frag_ids.insert(frag_ids.end(), buf.begin() + head, buf.begin() + tail);
            frag_off.push_back((int)frag_ids.size());
            frag_maxlen = std::max(frag_maxlen, (unsigned)n);
        }
    }
}

// form_curvelet_process.hpp fit_polyarcs_to_all_edgel_chains /
// fit_polyarc_to_chain and polyarcmap.hpp
// compute_best_fit_according_to_curvelets (CC_curve_model_3d.hpp
// update_best_fit, curvelet.hpp update_length)
void Former::stage_polyarcs() {
    const size_t F = frag_off.size() - 1;
    std::vector<std::vector<Arc>> part(F);
    std::vector<std::vector<int>> part_ids(F);
    std::vector<std::vector<double>> scratch(std::max(1, nthreads_));
    parallel_blocks(F, 8, [&](int t, size_t, size_t lo, size_t hi) {
        std::vector<double> &s = scratch[t];
        if (s.size() < 8 * (size_t)nc) s.resize(8 * (size_t)nc);
        for (size_t f = lo; f < hi; ++f) {
            const int *c = &frag_ids[frag_off[f]];
            const unsigned n = (unsigned)(frag_off[f + 1] - frag_off[f]);
            std::vector<Arc> &arcs_f = part[f];
            std::vector<int> &ids_f = part_ids[f];
            unsigned i = 0, j = 0;
            while (i < n - 1) {
                const int eA = c[i];
                int eB = c[i + 1];
                i++;
                BundleRef cm{&s[0], &s[nc]};
                if (dir_sign(eA, eB, sq_dist(pt[eA], pt[eB]), false) <= 0 || !bundle(eA, eB, cm)) continue;
                j = i;
                const size_t start = ids_f.size();
                ids_f.push_back(eA);
                ids_f.push_back(eB);
                int flip = 0;
                while (j < n - 1) {
                    eB = c[j + 1];
                    BundleRef cm2{&s[2 * nc], &s[3 * nc]};
                    if (dir_sign(eA, eB, sq_dist(pt[eA], pt[eB]), false) <= 0 || !bundle(eA, eB, cm2)) break;
                    double *nb = &s[(size_t)(4 + 2 * flip) * nc];
                    BundleRef cand{nb, nb + nc};
                    if (!intersect(cm, cm2, cand)) break;
                    cm = cand;
                    flip ^= 1;
                    ids_f.push_back(eB);
                    j++;
                }
                Arc arc;
                std::memset(&arc, 0, sizeof(arc));
                arc.n = (int)(ids_f.size() - start);
                arc.ids = (int)start;
                arc.info[0] = 1.0;
                best_fit(eA, cm, arc.info);
                properties(&ids_f[start], arc.n, eA, arc.info[8], arc.info[9]);
                arc.poly = false;
                // the end-point refit, applied min(F - f, 2) times
                const int reps = (F - f) >= 2 ? 2 : 1;
                const Pt p0 = pt[eA], p1 = pt[ids_f[start + arc.n / 2]], p2 = pt[ids_f[start + arc.n - 1]];
                double theta = arc.info[6], k = arc.info[7], cx = 0, cy = 0, angle = 0;
                for (int r = 0; r < reps; ++r) {
                    const Pt p = p0;
                    bool line = std::fabs(k) < kKth;
                    if (!line) {
                        const double d1 = sq_norm(p), d2 = sq_norm(p1), d3 = sq_norm(p2);
                        const double a11 = 2.0 * (p.x - p1.x), a12 = 2.0 * (p.y - p1.y),
                                     a21 = 2.0 * (p2.x - p1.x), a22 = 2.0 * (p2.y - p1.y);
                        const double b1 = d1 - d2, b2 = d3 - d2;
                        const double det = a11 * a22 - a21 * a12;
                        if (std::fabs(det) < kKth) {
                            line = true;
                        } else {
                            cx = (b1 * a22 - b2 * a12) / det;
                            cy = (b2 * a11 - b1 * a21) / det;
                            const Pt v1{p.x - cx, p.y - cy}, v2{p2.x - cx, p2.y - cy};
                            const double cr = v1.x * v2.y - v1.y * v2.x, dt_ = v1.x * v2.x + v1.y * v2.y;
                            angle = std::atan2(std::fabs(cr), dt_);
                            theta = std::atan2(cr * v1.x, -cr * v1.y);
                            k = 1.0 / std::sqrt(sq_norm(v2));
                            arc.poly = true;
                        }
                    }
                    if (line) {
                        k = 0.0;
                        theta = std::atan2(p2.y - p.y, p2.x - p.x);
                    }
                    if (std::fabs(angle) > 0 && std::fabs(k) > kKth && arc.poly) arc.info[8] = angle / std::fabs(k);
                }
                arc.info[4] = p0.x;
                arc.info[5] = p0.y;
                arc.info[6] = theta;
                arc.info[7] = k;
                if (arc.poly) {
                    arc.info[10] = cx;
                    arc.info[11] = cy;
                }
                arcs_f.push_back(arc);
                i = j;
            }
        }
    });
    arcs.clear();
    arc_ids.clear();
    arc_maxlen = 0;
    for (size_t f = 0; f < F; ++f) {
        const int base = (int)arc_ids.size();
        for (Arc a : part[f]) {
            a.ids += base;
            arc_maxlen = std::max(arc_maxlen, (unsigned)a.n);
            arcs.push_back(a);
        }
        arc_ids.insert(arc_ids.end(), part_ids[f].begin(), part_ids[f].end());
    }
}

void Former::execute(int nthreads) {
    nthreads_ = nthreads > 0 ? nthreads : (int)std::max(1u, std::thread::hardware_concurrency());
    precompute();
    stage_groups();
    stage_links_and_fragments();
    stage_polyarcs();
}

void Former::output(unsigned type, int *chain, double *info) const {
    unsigned h, w;
    output_size(type, h, w);
    const unsigned iw = type == 0 ? 10 : type == 1 ? 1 : 12;
    if (chain) std::memset(chain, 0, sizeof(int) * (size_t)h * w);
    if (info) std::memset(info, 0, sizeof(double) * (size_t)h * iw);
    if (type == 0) {
        for (unsigned r = 0; r < h; ++r) {
            const Curvelet &cv = cvlets[r];
            if (chain) {
                chain[r] = cv.anchor + 1;
                for (int q = 0; q < cv.n; ++q) chain[(size_t)(q + 1) * h + r] = cv_ids[cv.ids + q] + 1;
            }
            if (info)
                for (unsigned q = 0; q < 10; ++q) info[(size_t)q * h + r] = cv.info[q];
        }
    } else if (type == 1) {
        if (chain)
            for (unsigned r = 0; r < h; ++r)
                for (int q = frag_off[r]; q < frag_off[r + 1]; ++q)
                    chain[(size_t)(q - frag_off[r]) * h + r] = frag_ids[q] + 1;
    } else {
        for (unsigned r = 0; r < h; ++r) {
            const Arc &a = arcs[r];
            if (chain)
                for (int q = 0; q < a.n; ++q) chain[(size_t)q * h + r] = arc_ids[a.ids + q] + 1;
            if (info)
                for (unsigned q = 0; q < 12; ++q) info[(size_t)q * h + r] = a.info[q];
        }
    }
}

}  // namespace

struct toed_curvelets {
    Former *f;
    bool executed;
};

extern "C" {

int toed_curvelet_create(const double *edgeinfo, unsigned n_edgels, unsigned nrows, unsigned ncols, double rad,
                         double gap, double dx, double dt, double token_len, double max_k,
                         unsigned max_size_to_group, int centered, int bidirectional, toed_curvelets **out) {
    if (!out) return TOED_ERR_ARG;
    *out = nullptr;
    if ((n_edgels && !edgeinfo) || nrows == 0 || ncols == 0) return TOED_ERR_ARG;
    if (!(rad >= 0) || !(gap >= 0) || !(dx > 0) || !(dt > 0) || !(token_len > 0) || !(max_k >= 0)) return TOED_ERR_ARG;
    if (rad > 64 || dx / kStepX > 64 || dt / kStepT > 64 || max_size_to_group > 4096) return TOED_ERR_ARG;
    if ((unsigned long long)nrows * ncols > (1ull << 31) || n_edgels > (1u << 30)) return TOED_ERR_ARG;
    Params p{rad, gap, dx, dt, token_len, max_k, max_size_to_group, centered != 0, bidirectional != 0};
    toed_curvelets *h = new (std::nothrow) toed_curvelets;
    if (!h) return TOED_ERR_ALLOC;
    try {
        h->f = new Former(edgeinfo, n_edgels, nrows, ncols, p);
    } catch (...) {
        delete h;
        return TOED_ERR_ALLOC;
    }
    h->executed = false;
    *out = h;
    return TOED_OK;
}

int toed_curvelet_execute(toed_curvelets *h, int nthreads) {
    if (!h) return TOED_ERR_ARG;
    try {
        h->f->execute(nthreads);
    } catch (...) {
        return TOED_ERR_ALLOC;
    }
    h->executed = true;
    return TOED_OK;
}

int toed_curvelet_output_size(const toed_curvelets *h, unsigned output_type, unsigned *out_h, unsigned *out_w,
                              unsigned *info_w) {
    if (!h || !h->executed || output_type > 2 || !out_h || !out_w) return TOED_ERR_ARG;
    h->f->output_size(output_type, *out_h, *out_w);
    if (info_w) *info_w = output_type == 0 ? 10u : output_type == 1 ? 1u : 12u;
    return TOED_OK;
}

int toed_curvelet_output(const toed_curvelets *h, unsigned output_type, int *chain, double *info) {
    if (!h || !h->executed || output_type > 2) return TOED_ERR_ARG;
    h->f->output(output_type, chain, info);
    return TOED_OK;
}

int toed_curvelet_destroy(toed_curvelets *h) {
    if (!h) return TOED_ERR_ARG;
    delete h->f;
    delete h;
    return TOED_OK;
}

}  // extern "C"
