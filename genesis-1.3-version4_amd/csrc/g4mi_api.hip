// C-ABI of the g4mi library: device context and the step entry points.
//
// The host side here mirrors the reference's step sequencing
// (Gencore::run, src/Core/Gencore.cpp:102-160): beam track, field track,
// slippage, diagnostics.  All particle- and cell-sized work is done by the
// kernels in g4mi_kernels.hip on the context's stream; only per-step scalars
// are prepared on the host.
#include <algorithm>
#include <complex>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <map>
#include <mutex>
#include <tuple>

#include "g4mi_host.h"
#include "g4mi_kernels.h"
#include "g4mi_fft.h"
#include "g4mi_sort.h"
#include "g4mi_fno_api.h"
#include "g4mi_mixed.h"
#include "g4mi_load.h"

namespace g4mi {

thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int precision = G4MI_FP64;
    int ns = 0, ng = 0;
    double gridmax = 0, dgrid = 0, xks = 0, xlambda = 0, lambda0 = 0, sample = 1, slicelength = 0;
    int timerun = 1, periodic = 0;
    std::vector<long long> off;
    long long ntot = 0, maxn = 0;
    // device arrays
    long long *d_off = nullptr;
    double *d_g = nullptr, *d_th = nullptr, *d_x = nullptr, *d_y = nullptr, *d_px = nullptr,
           *d_py = nullptr;
    double *d_cur = nullptr, *d_esc = nullptr, *d_ez = nullptr, *d_mom = nullptr,
           *d_fsize = nullptr, *d_cen = nullptr, *d_rbound = nullptr, *d_dr = nullptr,
           *d_rmax = nullptr, *d_efield = nullptr;
    double2 *d_field = nullptr, *d_work = nullptr, *d_src = nullptr;
    int4 *d_box = nullptr;
    double2 *d_fac = nullptr;  // c | cbet | cwet, ngrid each
    int nblk = 0, blk_cap = 0;
    int *d_blk_slice = nullptr;
    long long *d_blk_begin = nullptr;
    double *d_diag_b = nullptr, *d_diag_f = nullptr;
    double *d_bunch = nullptr;  // harmonic bunching sums, [ns][8][2]
    unsigned long long *d_count = nullptr;
    int *d_atar = nullptr;
    // state
    int first = 0;
    double accuslip = 0;
    bool has_beam = false, has_field = false;
    // space charge
    int longrange = 0, sc_nz = 0, sc_nphi = 0, sc_ng = 100;
    double sc_rmax = 0;
    // ADI factor cache and the kernels eligible for it
    double adi_delz = 0, cstep = 0;
    int adi_tile = 0;  // 8 or 16: window of the fused recurrence kernel; 0: not eligible
    int adi_band_m = 0;  // 7, 8, 12 or 16: half width of the banded-inverse kernel; 0: not eligible
    int adi_klo = 0, adi_khi = -1;  // rows sharing the central band row
    double2 *d_band = nullptr;
    double *d_tilediag = nullptr;
    // field diagnostics sums written by the fused ADI of the last field
    // step, and the storage slices slippage has zeroed since then
    bool field_diag_fresh = false;
    std::vector<int> zeroed_since_adi;
    // ADI kernel chosen by the autotuner (-1: not tuned yet) and its timings:
    // 0 two-kernel Thomas, 1 fused recurrence, 2 fused recurrence with
    // prefetch, 3 banded inverse
    int adi_kernel = -1;
    bool adi_fixed = false;               // set by g4mi_set_schedule("adi_kernel", k >= 0)
    std::map<double, int> adi_choice;     // the tuner's kernel per step size
    float adi_ms[5] = {-1.f, -1.f, -1.f, -1.f, -1.f};
    // spectral solver (FieldSolverFFT) instead of the ADI, with the source
    // filter's settings; the propagator matrix is cached per delz
    int fft = 0, fft_filter = 0;
    int fft_variant = -1;  // complex GEMM variant chosen by the tuner (-1: not tuned yet)
    float fft_ms[kCgemmVariants] = {-1.f, -1.f, -1.f, -1.f};
    // propagation: 0 the two GEMMs with M, 1 FFT convolution (-1: the tuner
    // picks the faster); the convolution's length and tables per delz
    int fft_prop = -1, fft_conv_L = 0;
    float fft_conv_ms = -1.f;
    double2 *d_fft_hk = nullptr, *d_fft_tw = nullptr;
    // Bluestein DFT tables (forward, inverse): [L] kernel transform, then [ng] chirp
    double2 *d_fft_bf = nullptr, *d_fft_bi = nullptr;
    double fft_xc = 1, fft_yc = 1, fft_sig = 1;
    bool fft_ready = false;
    double fft_delz = 0;
    double2 *d_fft_m = nullptr, *d_fft_w = nullptr, *d_fft_wc = nullptr, *d_fft_u = nullptr,
            *d_fft_v = nullptr;
    double *d_fft_sig = nullptr, *d_fft_mom = nullptr, *d_fft_mompart = nullptr;
    // incoherent synchrotron radiation (Incoherent): the host's uniform
    // stream, one kick per beamlet staged through pinned memory
    Ran2 *incoh = nullptr;
    int incoh_loss = 0, incoh_spread = 0, incoh_nbins = 1;
    long long incoh_nb = 0;                 // beamlets in the whole beam
    long long *d_boff = nullptr;            // [ns+1] first beamlet of each slice
    double *d_bdg = nullptr, *h_bdg = nullptr;
    hipEvent_t incoh_ev = nullptr;          // the last kick upload has been read
    // wake loss per slice [eV/m] (Collective: wakeext + wakeint) or unset
    double *d_wake = nullptr, *d_wake_dg = nullptr, *d_wake_buf = nullptr;
    std::vector<double> h_wake, wake_applied;
    // fields at harmonics of the fundamental (FP64 form): own grid and
    // wavenumber, the fundamental's ring offset, Thomas ADI per field
    struct HField {
        int harm = 1, ng = 0;
        double gridmax = 0, dgrid = 0, xks = 0;
        double2 *d_field = nullptr, *d_work = nullptr, *d_src = nullptr, *d_fac = nullptr;
        int4 *d_box = nullptr;
        double fac_delz = 0, cstep = 0;
        bool fac_ready = false;
    };
    std::vector<HField> hf;
    // records of g4mi_run waiting for the host: their per-slice sums sit in
    // pinned slots (16 ns doubles each) filled by asynchronous copies
    struct PendingRecord {
        g4mi_diag_out out;
        int first;
        bool fresh;
        std::vector<int> zeroed;
        std::vector<double> wake;  // Beam::eloss at the record (auxiliary block)
    };
    double *h_recs = nullptr;
    double *d_recs = nullptr;  // h_recs' device address
    int rec_cap = 0;
    std::vector<PendingRecord> pend;
    // one4one beams: Beam::sort at sort markers, with its work buffers
    bool one4one = false;
    // the step's particle passes as two fused kernels (schedule "fp64_fused":
    // 1 on, 0 off, -1 on when the window has at least kFuseMinSlices slices,
    // since the fused kernels run one workgroup per slice while the separate
    // ones spread a slice's particles over blocks of 256)
    int fuse64 = -1;
    double *d_beamsum = nullptr;           // beam sums of the last fused step [8 ns]
    int4 *d_pbox = nullptr;                // in-grid footprint per slice after the first half step
    double *d_pmaxf = nullptr;             // per slice: largest sqrt(faw2) / gamma in the grid (deposit scale)
    bool beamsum_fresh = false;            // d_beamsum describes the current particles
    SortWork sw{};
    long long *d_off_new = nullptr;
    double wake_dz = 0;
    // timing: kEv events per step, recorded into a pool and folded lazily
    static constexpr int kEv = 9;
    static constexpr int kPoolSteps = 512;
    bool timing = false;
    std::vector<hipEvent_t> pool;
    int pool_used = 0;  // steps recorded in the pool
    double t_phase[G4MI_NTIMERS] = {};
    long long n_timed = 0;
    hipEvent_t *ev = nullptr;  // events of the step being recorded
    std::vector<void *> allocs;
    // ---- mixed form (precision G4MI_FP32)
    double gref = 0;                       // reference energy of the FP32 offsets
    float *m_dg = nullptr, *m_dth = nullptr, *m_x = nullptr, *m_y = nullptr, *m_px = nullptr,
          *m_py = nullptr;
    float *m_xs = nullptr, *m_ys = nullptr;  // post-t2 positions (the main loop's next front)
    double *m_base = nullptr;              // [ns] FP64 phase base
    int4 *m_pbox = nullptr;                // [ns] footprint after the first half step
    float2 *m_field[2] = {nullptr, nullptr}, *m_src = nullptr;
    double *m_beamsum = nullptr, *m_egsum = nullptr, *m_tilediag = nullptr, *m_rec = nullptr;
    double *m_mom_next = nullptr;          // [ns] second slice-size buffer (steps alternate with d_fsize)
    int m_nrec = 0;
    MixStep *m_steps = nullptr;            // device step table
    int m_steps_cap = 0;
    MixState *m_st = nullptr;
    float2 *m_bands = nullptr;             // device band table
    MixBand *m_binfo = nullptr;
    int m_binfo_cap = 0;
    std::vector<float2> h_bands;           // host copy of the band table
    std::vector<MixBand> h_binfo;
    std::map<double, int> band_of_dz;      // step size -> band index
    std::vector<int> band_tmpl;            // template half width of each band
    bool m_refine = true;                  // FP64 residual correction in the ADI (schedule "adi_refine")
    bool m_graphs = true;                  // replay captured step graphs
    int m_adi_mode = kAdiWide;             // mixed ADI kernel (schedule "mixed_adi"): tile 0, strips 1, wide 2
    int m_main_nt = 384;                   // threads per slice of the main kernel ("main_nt")
    int m_main_lds = 0;                    // its LDS region: 0 = 68 KiB, 1 = 44 KiB, 2 = 32 KiB ("main_lds")
    bool m_persistent = false;             // the whole step as one cooperative launch ("persistent")
    bool m_fuse_front = true;              // main kernels run the next step's front ("mixed_fuse_front")
    bool m_autotune = true;                // time the schedules on the first steps of a run
    bool m_tuned = false;
    std::string m_tune_log;                // mean ms per candidate, for the schedule report
    int m_tune_i = 0;                      // tuning steps done
    double m_tune_main[8] = {0, 0, 0, 0, 0, 0, 0, 0}, m_tune_adi[3] = {0, 0, 0}, m_tune_graph = -1;
    int m_tune_nmain[8] = {0, 0, 0, 0, 0, 0, 0, 0}, m_tune_nadi[3] = {0, 0, 0};
    hipEvent_t m_tune_ev[9] = {};
    std::map<std::tuple<int, int, int, int, int, int, int, int, int>, hipGraphExec_t> m_graph;
    double *m_rec_pin = nullptr;           // pinned host mirror of the record ring
    hipEvent_t m_rec_ev[2] = {nullptr, nullptr};  // copy of each half of the ring done
    std::vector<double> h_cur;             // currents [ns] (host copy)
    double *d_aux = nullptr;               // auxiliary diagnostics [16 ns]
    // learned field step (mixed form): 0 detached, 1 capture, 2 replace the ADI
    Fno *fno = nullptr;
    int fno_mode = 0, fno_every = 1;
    double fno_thr = 1e-3;
    long long fno_steps = 0, fno_adi_steps = 0, fno_nres_adi = 0;
    double fno_res_sum = 0, fno_res_adi_sum = 0;
    // judge steps (every fno_every-th step of replace mode): the network's
    // step is kept in m_judge and compared with the ADI's
    float2 *m_judge = nullptr;
    double *d_judge = nullptr;
    long long fno_njudge = 0, fno_nres_all = 0;
    // replace mode backs off after kFnoRejectRun rejections in a row: the
    // next kFnoSkip steps run the ADI without inference
    int fno_reject_run = 0, fno_skip = 0;
    long long fno_skipped = 0;
    double fno_judge_field = 0, fno_judge_update = 0, fno_res_all_sum = 0;
};

template <class T>
static int dalloc(Ctx *c, T **p, size_t n) {
    if (n == 0) n = 1;
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, n * sizeof(T));
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
        return G4MI_E_HIP;
    }
    c->allocs.push_back(q);
    *p = (T *)q;
    return G4MI_OK;
}


static TransParams transverse_params(const Ctx *c, const g4mi_step_params &st, double gamma0,
                                     double delz, bool last) {
    TransParams p{};
    const bool inund = st.aw > 0;  // Undulator::getUndulatorParameters
    const double aw = inund ? st.aw : 0, dax = inund ? st.ax : 0, day = inund ? st.ay : 0;
    const double kx = inund ? st.kx : 0, ky = inund ? st.ky : 0;
    const double qf = st.qf;
    const double dqx = qf == 0 ? 0 : st.qx, dqy = qf == 0 ? 0 : st.qy;
    const double betpar0 = sqrt(1 - (1 + aw * aw) / gamma0 / gamma0);
    const double qquad = qf * gamma0;
    const double qnatx = kx * aw * aw / gamma0 / betpar0;
    const double qnaty = ky * aw * aw / gamma0 / betpar0;
    p.aw = aw;
    p.delz = delz;
    p.qx = qquad + qnatx;
    p.qy = -qquad + qnaty;
    p.xoff = qquad * dqx + qnatx * dax;
    p.yoff = -qquad * dqy + qnaty * day;
    if (last) {
        if (st.cx != 0 || st.cy != 0) {
            p.corrector = 1;
            p.ccx = st.cx * gamma0;
            p.ccy = st.cy * gamma0;
        }
    } else if (st.chic_angle != 0) {
        double m[16];
        chicane_matrix(st.chic_angle, st.chic_lb, st.chic_ld, st.chic_lt, m);
        p.chicane = 1;
        p.m00 = m[0];
        p.m01 = m[1];
        p.m10 = m[4];
        p.m11 = m[5];
        p.m22 = m[10];
        p.m23 = m[11];
        p.m32 = m[14];
        p.m33 = m[15];
    }
    if (p.qx != 0) {
        p.xoff = p.xoff / p.qx;
        p.kindx = p.qx > 0 ? 1 : -1;
    }
    if (p.qy != 0) {
        p.yoff = p.yoff / p.qy;
        p.kindy = p.qy > 0 ? 1 : -1;
    }
    (void)c;
    return p;
}

static void build_block_table(const std::vector<long long> &off, std::vector<int> &bs,
                              std::vector<long long> &bb) {
    bs.clear();
    bb.clear();
    for (size_t is = 0; is + 1 < off.size(); is++)
        for (long long a = off[is]; a < off[is + 1]; a += 256) {
            bs.push_back((int)is);
            bb.push_back(a);
        }
}

static int ensure_factors(Ctx *c, double delz) {
    if (delz == c->adi_delz) return G4MI_OK;  // FieldSolverADI::init caching on delz
    c->adi_delz = delz;
    std::vector<double> h(6 * (size_t)c->ng);
    c->cstep = adi_factors(c->ng, delz, c->dgrid, c->xks, h.data(), h.data() + 2 * c->ng,
                           h.data() + 4 * c->ng);
    // Window of the fused kernel: a back-substitution started T cells early
    // is off by at most max|cwet|^(T+1) relative; accept it below 2^-66.
    double wmax = 0;
    for (int i = 0; i < c->ng; i++)
        wmax = std::max(wmax, std::hypot(h[4 * c->ng + 2 * i], h[4 * c->ng + 2 * i + 1]));
    const double bound = std::ldexp(1.0, -66);
    c->adi_tile = 0;
    if (c->ng <= 256) {
        if (std::pow(wmax, 9) < bound) c->adi_tile = 8;
        else if (std::pow(wmax, 17) < bound) c->adi_tile = 16;
    }
    // banded inverse: accept the narrowest band whose dropped entries are
    // all below 2^-60 of the diagonal (1/128 ulp of the result)
    c->adi_band_m = 0;
    for (int m : {7, 8, 12, 16}) {
        std::vector<double> band(2 * (size_t)c->ng * (2 * m + 1));
        const double worst = adi_band(c->ng, h.data(), h.data() + 2 * c->ng,
                                      h.data() + 4 * c->ng, m, band.data());
        if (worst < std::ldexp(1.0, -60)) {
            // interior rows whose band equals the central row's to 2^-60
            const int W = 2 * m + 1, kc = c->ng / 2;
            auto same = [&](int k) {
                double d = 0, dia = std::hypot(band[2 * ((size_t)kc * W + m)], band[2 * ((size_t)kc * W + m) + 1]);
                for (int q = 0; q < W; q++)
                    d = std::max(d, std::hypot(band[2 * ((size_t)k * W + q)] - band[2 * ((size_t)kc * W + q)],
                                               band[2 * ((size_t)k * W + q) + 1] - band[2 * ((size_t)kc * W + q) + 1]));
                return d <= std::ldexp(1.0, -60) * dia;
            };
            int lo = kc, hi = kc;
            while (lo > 0 && same(lo - 1)) lo--;
            while (hi < c->ng - 1 && same(hi + 1)) hi++;
            // the kernel takes the band of row (klo+khi)/2: keep the range centred
            const int half = std::min(kc - lo, hi - kc);
            c->adi_klo = kc - half;
            c->adi_khi = kc + half;
            if (c->d_band) (void)hipFree(c->d_band);
            c->d_band = nullptr;
            G4_HIP(hipMalloc(&c->d_band, band.size() * sizeof(double)));
            G4_HIP(hipMemcpy(c->d_band, band.data(), band.size() * sizeof(double),
                             hipMemcpyHostToDevice));
            if (!c->d_tilediag)
                G4_HIP(hipMalloc(&c->d_tilediag,
                                 (size_t)c->ns * adi_band_tiles(c->ng) * 8 * sizeof(double)));
            c->adi_band_m = m;
            break;
        }
    }
    if (!c->adi_fixed) {
        // the tuner's choice for this step size, or tune it at the next field step
        auto it = c->adi_choice.find(delz);
        c->adi_kernel = it != c->adi_choice.end() ? it->second : -1;
    } else if (c->adi_kernel == 1 || c->adi_kernel == 2) {
        if (!c->adi_tile) c->adi_kernel = -1;      // re-tune if no longer eligible
    } else if (c->adi_kernel == 3 && !c->adi_band_m) {
        c->adi_kernel = -1;
    } else if (c->adi_kernel == 4 && !adi_wide64_ok(c->ng, c->adi_band_m, c->adi_klo, c->adi_khi)) {
        c->adi_kernel = -1;
    }
    G4_HIP(hipMemcpyAsync(c->d_fac, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                          c->stream));
    G4_HIP(hipStreamSynchronize(c->stream));
    return G4MI_OK;
}

// Incoherent::apply (src/Core/Incoherent.cpp:35-82) and Collective::apply
// (src/Core/Collective.cpp:107-128) after the push.  The incoherent kicks
// of one step are drawn on the host from the run's single uniform stream,
// one per beamlet in slice and particle order, so the device adds exactly
// the reference's values; the wake kick is eloss * delz / eev per slice.
static int energy_kicks(Ctx *c, const g4mi_step_params &st, double gamma0, double delz) {
    const double *bdg = nullptr, *sdg = nullptr;
    if (c->incoh && st.aw > 0 && (c->incoh_loss || c->incoh_spread)) {
        const double gam0 = gamma0, awz = st.aw, xkw0 = st.ku;
        double dgamsig = 1.015e-27 * xkw0 * xkw0 * awz * awz;
        if (st.helical)
            dgamsig *= 1.42 * awz + 1. / (1. + 1.5 * awz + 0.95 * awz * awz);
        else
            dgamsig *= 1.697 * awz + 1. / (1. + 1.88 * awz + 0.8 * awz * awz);
        if (!c->incoh_spread) dgamsig = 0;
        dgamsig = sqrt(dgamsig * gam0 * gam0 * gam0 * gam0 * xkw0 * delz) * sqrt(3.);
        double dgamavg = xkw0 * gam0 * awz;
        if (!c->incoh_loss) dgamavg = 0;
        dgamavg = dgamavg * dgamavg * 1.88e-15 * delz;
        G4_HIP(hipEventSynchronize(c->incoh_ev));  // the previous upload has been read
        for (long long b = 0; b < c->incoh_nb; b++) c->h_bdg[b] = -dgamavg + dgamsig * (2 * c->incoh->next() - 1);
        G4_HIP(hipMemcpyAsync(c->d_bdg, c->h_bdg, c->incoh_nb * sizeof(double), hipMemcpyHostToDevice, c->stream));
        G4_HIP(hipEventRecord(c->incoh_ev, c->stream));
        bdg = c->d_bdg;
    }
    if (c->d_wake) {
        if (delz != c->wake_dz) {
            std::vector<double> dg(c->ns);
            for (int s = 0; s < c->ns; s++) dg[s] = c->h_wake[s] * delz / kEev;
            G4_HIP(hipMemcpyAsync(c->d_wake_dg, dg.data(), dg.size() * sizeof(double), hipMemcpyHostToDevice,
                                  c->stream));
            G4_HIP(hipStreamSynchronize(c->stream));
            c->wake_dz = delz;
        }
        sdg = c->d_wake_dg;
        c->wake_applied = c->h_wake;
    }
    launch_kicks(c->stream, c->ns, c->d_off, c->d_g, bdg, c->d_boff, c->incoh_nbins, sdg);
    return G4MI_OK;
}

static int beam_track(Ctx *c, const g4mi_step_params &st, double gamma0, double delz) {
    if (!c->has_beam) {
        set_error("beam not uploaded");
        return G4MI_E_STATE;
    }
    TransParams t1 = transverse_params(c, st, gamma0, delz * 0.5, false);
    if (c->ev) G4_HIP(hipEventRecord(c->ev[0], c->stream));
    launch_transverse(c->stream, c->ntot, c->d_g, c->d_x, c->d_y, c->d_px, c->d_py, t1);

    // BeamSolver::advance
    if (c->ev) G4_HIP(hipEventRecord(c->ev[1], c->stream));
    const double aw = st.aw;
    const bool sc = (c->sc_nz > 0) && (c->sc_ng > 2);
    if (c->longrange || sc) {
        launch_slice_moments(c->stream, c->ns, c->d_off, c->d_x, c->d_y, c->d_mom);
        launch_slice_sizes(c->stream, c->ns, c->d_off, c->d_cur, c->d_mom, c->d_fsize, c->d_cen);
    }
    if (c->longrange) {
        const double gamma = gamma0 / sqrt(1 + aw * aw);
        const double scl = c->slicelength / 2. / asin(1) / 299792458.0 / 2 / 8.85e-12;
        launch_lsc_long(c->stream, c->ns, c->d_cur, c->d_fsize, c->slicelength, gamma, scl,
                        c->d_esc);
    }
    if (sc) {
        launch_sc_rbound(c->stream, c->ns, c->d_off, c->d_x, c->d_y, c->d_cen, c->d_rbound);
        launch_sc_scan(c->stream, c->ns, c->d_off, c->d_rbound, c->d_rmax, c->sc_ng, c->d_dr);
        const double gz2 = gamma0 * gamma0 / (1 + aw * aw);
        const double ks = 4 * asin(1) / c->lambda0;
        launch_sc_solve(c->stream, c->ns, c->d_off, c->d_x, c->d_y, c->d_th, c->d_cur, c->d_cen,
                        c->d_dr, c->sc_ng, c->sc_nz, c->sc_nphi, ks, gz2, c->d_ez, c->d_efield);
    }
    if (c->ev) G4_HIP(hipEventRecord(c->ev[2], c->stream));
    PushParams p{};
    p.delz = delz;
    p.xks = c->xks;
    p.rtmp = coupling(st, 1) / c->xks;
    p.xku = st.ku;
    if (p.xku == 0) p.xku = c->xks * 0.5 / gamma0 / gamma0;
    p.aw = aw;
    p.autophase = st.phaseshift;
    p.ax = st.ax;
    p.ay = st.ay;
    p.kx = st.kx;
    p.ky = st.ky;
    p.gradx = st.gradx;
    p.grady = st.grady;
    p.gridmax = c->gridmax;
    p.dgrid = c->dgrid;
    p.ngrid = c->ng;
    p.first = c->first;
    p.nslice = c->ns;
    p.use_sc = sc ? 1 : 0;
    p.nh = (int)c->hf.size();
    for (int k = 0; k < p.nh; k++) {
        const Ctx::HField &h = c->hf[k];
        p.hg[k].field = h.d_field;
        p.hg[k].gridmax = h.gridmax;
        p.hg[k].dgrid = h.dgrid;
        p.hg[k].ngrid = h.ng;
        p.hg[k].harm = (double)h.harm;
        p.hg[k].rtmp = coupling(st, h.harm) / h.xks;  // BeamSolver.cpp:28
    }
    launch_push(c->stream, c->nblk, c->d_blk_slice, c->d_blk_begin, c->d_off, c->d_g, c->d_th,
                c->d_x, c->d_y, c->d_px, c->d_py, c->d_field, c->d_esc, c->d_ez, p);
    if (c->ev) G4_HIP(hipEventRecord(c->ev[3], c->stream));
    G4_TRY(energy_kicks(c, st, gamma0, delz));
    if (st.chic_angle != 0) {
        const double r56 = chicane_r56(st.chic_angle, st.chic_lb, st.chic_ld, c->lambda0, gamma0);
        launch_r56(c->stream, c->ntot, c->d_g, c->d_th, r56, gamma0);
    }
    TransParams t2 = transverse_params(c, st, gamma0, delz * 0.5, true);
    launch_transverse(c->stream, c->ntot, c->d_g, c->d_x, c->d_y, c->d_px, c->d_py, t2);
    if (c->ev) G4_HIP(hipEventRecord(c->ev[4], c->stream));
    G4_HIP(hipGetLastError());
    return G4MI_OK;
}

// One ADI launch of kernel `kind` on field buffer `fin` (the banded kernel
// writes `fout`, the others work in place using `fout` as scratch).
static void adi_launch(Ctx *c, int kind, const AdiFactors &f, double2 *fin, double2 *fout) {
    if (kind == 4)
        (void)launch_adi_wide64(c->stream, c->ns, fin, fout, c->d_src, c->d_box, c->d_band, c->cstep, c->ng,
                                c->first, c->adi_band_m, c->adi_klo, c->adi_khi, c->d_diag_f);
    else if (kind == 3)
        launch_adi_band(c->stream, c->ns, fin, fout, c->d_src, c->d_box, c->d_band, c->cstep,
                        c->ng, c->first, c->adi_band_m, c->adi_klo, c->adi_khi, c->d_tilediag,
                        c->d_diag_f);
    else if (kind == 1 || kind == 2)
        launch_adi_fused(c->stream, c->ns, fin, fout, c->d_src, c->d_box, f, c->first,
                         c->d_diag_f, c->adi_tile, kind == 2);
    else
        launch_adi(c->stream, c->ns, fin, fout, c->d_src, c->d_box, f, c->first);
}

// Autotuner of the ADI: run each eligible kernel on a scratch copy of the
// current field (same sources and boxes) and keep the fastest.  Runs
// before the first field step of a context and after a step-size change
// that alters eligibility.
static int tune_adi(Ctx *c, const AdiFactors &f) {
    const size_t bytes = (size_t)c->ns * c->ng * c->ng * sizeof(double2);
    double2 *scratch = nullptr;
    G4_HIP(hipMalloc(&scratch, bytes));
    hipEvent_t e0, e1;
    G4_HIP(hipEventCreate(&e0));
    G4_HIP(hipEventCreate(&e1));
    G4_HIP(hipMemcpyAsync(scratch, c->d_field, bytes, hipMemcpyDeviceToDevice, c->stream));
    int best_kind = 0;
    float best_ms = 1e30f;
    for (int kind = 0; kind < 5; kind++) {
        c->adi_ms[kind] = -1.f;
        if ((kind == 1 || kind == 2) && !c->adi_tile) continue;
        if (kind == 3 && !c->adi_band_m) continue;
        if (kind == 4 && !adi_wide64_ok(c->ng, c->adi_band_m, c->adi_klo, c->adi_khi)) continue;
        adi_launch(c, kind, f, scratch, c->d_work);  // warm-up
        float best = 1e30f;
        for (int rep = 0; rep < 3; rep++) {
            G4_HIP(hipEventRecord(e0, c->stream));
            adi_launch(c, kind, f, scratch, c->d_work);
            G4_HIP(hipEventRecord(e1, c->stream));
            G4_HIP(hipEventSynchronize(e1));
            float ms = 0;
            G4_HIP(hipEventElapsedTime(&ms, e0, e1));
            best = std::min(best, ms);
        }
        c->adi_ms[kind] = best;
        if (best < best_ms) {
            best_ms = best;
            best_kind = kind;
        }
    }
    G4_HIP(hipGetLastError());
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    G4_HIP(hipFree(scratch));
    c->adi_kernel = best_kind;
    c->adi_choice[c->adi_delz] = best_kind;
    return G4MI_OK;
}

// FieldSolverFFT::FFT for every slice (src/Core/FieldSolverFFT.cpp:62-119)
// after the deposit: E' = M E M^T + 2 S, or with the source filter
// E' = M E M^T + 2 IDFT(sigmoid DFT(S)) / n^2, as complex FP64 GEMMs.
// Tuner of the complex GEMM: each variant runs the row product E M^T of the
// current field into the work buffer (which the step overwrites anyway)
// and the fastest is kept for both products of every later step.
static int tune_fft(Ctx *c) {
    hipEvent_t e0, e1;
    G4_HIP(hipEventCreate(&e0));
    G4_HIP(hipEventCreate(&e1));
    CgemmArgs g{};
    g.ng = c->ng;
    g.ns = c->ns;
    g.nrows = (long long)c->ns * c->ng;
    g.scale = 1.0;
    g.a = c->d_field;
    g.mat = c->d_fft_m;
    g.c = c->d_work;
    int best = c->fft_variant;
    float best_ms = 1e30f;
    for (int v = 0; v < kCgemmVariants; v++) {
        if (c->fft_variant >= 0 && v != c->fft_variant) continue;  // a fixed variant is only timed
        g.variant = v;
        G4_TRY(launch_cgemm_rows(c->stream, g));  // warm-up
        float b = 1e30f;
        for (int rep = 0; rep < 3; rep++) {
            G4_HIP(hipEventRecord(e0, c->stream));
            G4_TRY(launch_cgemm_rows(c->stream, g));
            G4_HIP(hipEventRecord(e1, c->stream));
            G4_HIP(hipEventSynchronize(e1));
            float ms = 0;
            G4_HIP(hipEventElapsedTime(&ms, e0, e1));
            b = std::min(b, ms);
        }
        c->fft_ms[v] = b;
        if (b < best_ms || best < 0) {
            best_ms = b;
            best = v;
        }
    }
    c->fft_variant = best;
    if (c->fft_prop < 0) {  // the convolution's rows pass against the best GEMM's
        c->fft_prop = 0;
        c->fft_conv_ms = -1.f;
        if (c->fft_conv_L) {
            ConvArgs a{};
            a.in = c->d_field;
            a.out = c->d_work;
            a.hk = c->d_fft_hk;
            a.tw = c->d_fft_tw;
            a.nrows = (long long)c->ns * c->ng;
            a.ng = c->ng;
            a.ns = c->ns;
            a.L = c->fft_conv_L;
            G4_TRY(launch_fft_conv_rows(c->stream, a));  // warm-up
            float b = 1e30f;
            for (int rep = 0; rep < 3; rep++) {
                G4_HIP(hipEventRecord(e0, c->stream));
                G4_TRY(launch_fft_conv_rows(c->stream, a));
                G4_HIP(hipEventRecord(e1, c->stream));
                G4_HIP(hipEventSynchronize(e1));
                float ms = 0;
                G4_HIP(hipEventElapsedTime(&ms, e0, e1));
                b = std::min(b, ms);
            }
            c->fft_conv_ms = b;
            if (b < c->fft_ms[best]) c->fft_prop = 1;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return G4MI_OK;
}

static int ensure_bluestein(Ctx *c);
static ConvArgs bluestein_pass(const Ctx *c, int sign, const double2 *in, double2 *out);

static int fft_field_step(Ctx *c, double delz) {
    const int ng = c->ng;
    const size_t nn = (size_t)ng * ng;
    if (!c->d_fft_m) G4_TRY(dalloc(c, &c->d_fft_m, nn));
    if (!c->fft_ready || delz != c->fft_delz) {  // FieldSolverFFT::init keeps delz per call
        std::vector<double> m(2 * nn);
        fft_propagator(ng, c->dgrid, c->xks, delz, m.data());
        G4_HIP(hipMemcpyAsync(c->d_fft_m, m.data(), m.size() * sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
        G4_HIP(hipStreamSynchronize(c->stream));
        c->fft_conv_L = fft_conv_length(ng);
        if (c->fft_conv_L) {
            if (!c->d_fft_hk) {
                G4_TRY(dalloc(c, &c->d_fft_hk, 512));
                G4_TRY(dalloc(c, &c->d_fft_tw, 512));
            }
            std::vector<double> hk(2 * (size_t)c->fft_conv_L), tw(2 * (size_t)c->fft_conv_L);
            fft_conv_tables(ng, c->dgrid, c->xks, delz, hk.data(), tw.data());
            G4_HIP(hipMemcpyAsync(c->d_fft_hk, hk.data(), hk.size() * sizeof(double), hipMemcpyHostToDevice,
                                  c->stream));
            G4_HIP(hipMemcpyAsync(c->d_fft_tw, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice,
                                  c->stream));
            G4_HIP(hipStreamSynchronize(c->stream));
        }
        c->fft_delz = delz;
        c->fft_ready = true;
    }
    if (c->fft_prop == 1 && !c->fft_conv_L) {
        set_error("fft_propagation 1: the FFT convolution supports ngrid <= 256");
        return G4MI_E_ARG;
    }
    if (c->fft_variant < 0 || c->fft_prop < 0) G4_TRY(tune_fft(c));
    CgemmArgs g{};
    g.ng = ng;
    g.ns = c->ns;
    g.nrows = (long long)c->ns * ng;
    g.box = c->d_box;
    g.scale = 1.0;
    g.variant = c->fft_variant;
    if (c->fft_filter && c->fft_prop == 1) {  // the same four DFT passes by Bluestein
        G4_TRY(ensure_bluestein(c));
        ConvArgs a = bluestein_pass(c, -1, c->d_src, c->d_fft_u);
        a.mask = 1;
        a.box = c->d_box;
        G4_TRY(launch_fft_conv_rows(c->stream, a));
        a = bluestein_pass(c, -1, c->d_fft_u, c->d_fft_v);
        a.ptw = c->d_fft_sig;
        G4_TRY(launch_fft_conv_cols(c->stream, a));
        G4_TRY(launch_fft_conv_rows(c->stream, bluestein_pass(c, +1, c->d_fft_v, c->d_fft_u)));
        a = bluestein_pass(c, +1, c->d_fft_u, c->d_fft_v);
        a.scale = 1.0 / (double)nn;
        G4_TRY(launch_fft_conv_cols(c->stream, a));
    } else if (c->fft_filter) {
        g.a = c->d_src;  // DFT of the source planes: U = S W^T, V = sigmoid (W U)
        g.mask_a = 1;
        g.mat = c->d_fft_w;
        g.c = c->d_fft_u;
        G4_TRY(launch_cgemm_rows(c->stream, g));
        g.mask_a = 0;
        g.b = c->d_fft_u;
        g.ptw = c->d_fft_sig;
        g.c = c->d_fft_v;
        G4_TRY(launch_cgemm_cols(c->stream, g));
        g.ptw = nullptr;  // back transform: Y = W* (V W*^T) / n^2
        g.a = c->d_fft_v;
        g.mat = c->d_fft_wc;
        g.c = c->d_fft_u;
        G4_TRY(launch_cgemm_rows(c->stream, g));
        g.b = c->d_fft_u;
        g.c = c->d_fft_v;
        g.scale = 1.0 / (double)nn;
        G4_TRY(launch_cgemm_cols(c->stream, g));
        g.scale = 1.0;
    }
    if (c->fft_prop == 1) {  // the same two products as FFT convolutions of every line
        ConvArgs a{};
        a.hk = c->d_fft_hk;
        a.tw = c->d_fft_tw;
        a.nrows = (long long)c->ns * ng;
        a.ng = ng;
        a.ns = c->ns;
        a.L = c->fft_conv_L;
        a.in = c->d_field;
        a.out = c->d_work;
        G4_TRY(launch_fft_conv_rows(c->stream, a));
        a.in = c->d_work;
        a.out = c->d_field;
        a.box = c->d_box;
        a.ring = 1;
        a.first = c->first;
        if (c->fft_filter)
            a.add = c->d_fft_v;
        else
            a.src = c->d_src;
        return launch_fft_conv_cols(c->stream, a);
    }
    g.a = c->d_field;  // T = E M^T
    g.mat = c->d_fft_m;
    g.c = c->d_work;
    G4_TRY(launch_cgemm_rows(c->stream, g));
    g.b = c->d_work;  // E = M T + 2 (S or Y) of beam slice (s - first) mod Ns
    g.c = c->d_field;
    g.ring = 1;
    g.first = c->first;
    if (c->fft_filter)
        g.add = c->d_fft_v;
    else
        g.src = c->d_src;
    G4_TRY(launch_cgemm_cols(c->stream, g));
    return G4MI_OK;
}

// Field::track of every harmonic field (FieldSolverADI per field, Field.cpp
// :177-185): the source at harmonic h (odd h only, FieldSolverADI.cpp:19-22)
// and the two-kernel Thomas ADI with the field's own factors.
static int harmonics_track(Ctx *c, const g4mi_step_params &st, double delz) {
    for (Ctx::HField &h : c->hf) {
        if (!h.fac_ready || delz != h.fac_delz) {  // FieldSolverADI::init for this field
            std::vector<double> f(6 * (size_t)h.ng);
            h.cstep = adi_factors(h.ng, delz, h.dgrid, h.xks, f.data(), f.data() + 2 * h.ng, f.data() + 4 * h.ng);
            G4_HIP(hipMemcpyAsync(h.d_fac, f.data(), f.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
            G4_HIP(hipStreamSynchronize(c->stream));
            h.fac_delz = delz;
            h.fac_ready = true;
        }
        DepositParams d{};
        d.enabled = (st.aw > 0 && (h.harm % 2 == 1)) ? 1 : 0;
        d.harm = (double)h.harm;
        d.fcoup = coupling(st, h.harm);
        d.xks = h.xks;
        d.delz = delz;
        d.dgrid = h.dgrid;
        d.gridmax = h.gridmax;
        d.ax = st.ax;
        d.ay = st.ay;
        d.kx = st.kx;
        d.ky = st.ky;
        d.gradx = st.gradx;
        d.grady = st.grady;
        d.ngrid = h.ng;
        launch_deposit(c->stream, c->ns, c->d_off, c->d_g, c->d_th, c->d_x, c->d_y, c->d_cur, h.d_src, h.d_box, d);
        AdiFactors f{h.d_fac, h.d_fac + h.ng, h.d_fac + 2 * h.ng, h.cstep, h.ng};
        launch_adi(c->stream, c->ns, h.d_field, h.d_work, h.d_src, h.d_box, f, c->first);
    }
    G4_HIP(hipGetLastError());
    return G4MI_OK;
}

static int field_track(Ctx *c, const g4mi_step_params &st, double delz, bool deposit = true) {
    if (!c->has_field) {
        set_error("field not uploaded");
        return G4MI_E_STATE;
    }
    if (!c->hf.empty()) G4_TRY(harmonics_track(c, st, delz));
    if (!c->fft) G4_TRY(ensure_factors(c, delz));
    if (c->ev) G4_HIP(hipEventRecord(c->ev[5], c->stream));
    DepositParams d{};
    d.enabled = st.aw > 0 ? 1 : 0;  // in undulator, fundamental
    d.harm = 1.0;
    d.fcoup = coupling(st, 1);
    d.xks = c->xks;
    d.delz = delz;
    d.dgrid = c->dgrid;
    d.gridmax = c->gridmax;
    d.ax = st.ax;
    d.ay = st.ay;
    d.kx = st.kx;
    d.ky = st.ky;
    d.gradx = st.gradx;
    d.grady = st.grady;
    d.ngrid = c->ng;
    if (deposit)
        launch_deposit(c->stream, c->ns, c->d_off, c->d_g, c->d_th, c->d_x, c->d_y, c->d_cur,
                       c->d_src, c->d_box, d);
    if (c->ev) G4_HIP(hipEventRecord(c->ev[6], c->stream));
    if (c->fft) {
        G4_TRY(fft_field_step(c, delz));
        c->field_diag_fresh = false;
        c->zeroed_since_adi.clear();
        if (c->ev) G4_HIP(hipEventRecord(c->ev[7], c->stream));
        return G4MI_OK;
    }
    AdiFactors f{c->d_fac, c->d_fac + c->ng, c->d_fac + 2 * c->ng, c->cstep, c->ng};
    if (c->adi_kernel < 0) G4_TRY(tune_adi(c, f));
    adi_launch(c, c->adi_kernel, f, c->d_field, c->d_work);
    if (c->adi_kernel >= 3) std::swap(c->d_field, c->d_work);  // result is in the other buffer
    c->field_diag_fresh = c->adi_kernel != 0;
    c->zeroed_since_adi.clear();
    G4_HIP(hipGetLastError());
    if (c->ev) G4_HIP(hipEventRecord(c->ev[7], c->stream));
    return G4MI_OK;
}

static int slippage(Ctx *c, double slip) {
    if (!c->timerun) return G4MI_OK;  // Control::applySlippage :159
    c->accuslip += slip;
    int direction = 1;
    const long long cells = (long long)c->ng * c->ng;
    while (fabs(c->accuslip) > (c->sample * 0.8)) {
        if (c->accuslip < 0) direction = -1;
        c->accuslip -= c->sample * direction;
        int last = (c->first + c->ns - 1) % c->ns;
        if (direction < 0) last = (last + 1) % c->ns;
        if (!c->periodic) {
            launch_zero(c->stream, c->d_field + (long long)last * cells, cells);
            c->zeroed_since_adi.push_back(last);
            for (const Ctx::HField &h : c->hf)  // every field slips alike
                launch_zero(c->stream, h.d_field + (long long)last * h.ng * h.ng, (long long)h.ng * h.ng);
        }
        c->first = last;
        if (direction < 0) c->first = (last + 1) % c->ns;
    }
    return G4MI_OK;
}

// Normalisation of the per-slice sums: hb[8*is] beam sums per beam slice,
// hf[8*is0] field sums per storage slice, mapped to beam slices through the
// ring offset `first`.
// DiagField::getValues (Diagnostic.cpp:754-969) from the per-storage-slice
// field sums hf[8 is0] of a field on a grid of spacing dgrid with
// wavenumber ks; first = ring offset.
static void finish_field(const std::vector<double> &hf, int ns, int first, double dgrid, double ks,
                         double slicelength, g4mi_diag_out *o) {
    const double scl = dgrid * kEev / ks;
    double g_pow = 0, fx1 = 0, fx2 = 0, fy1 = 0, fy2 = 0, g_ff = 0, g_in = 0;
    for (int is0 = 0; is0 < ns; is0++) {
        const int is = (ns + is0 - first) % ns;
        const double *v = &hf[8 * is0];
        double power = v[0], x1 = v[1], x2 = v[2], y1 = v[3], y2 = v[4];
        g_pow += power;
        fx1 += x1;
        fx2 += x2;
        fy1 += y1;
        fy2 += y2;
        if (o->ff_phase) o->ff_phase[is] = atan2(v[6], v[5]);
        if (power > 0) {
            x1 /= power;
            x2 /= power;
            y1 /= power;
            y2 /= power;
        }
        x2 = sqrt(fabs(x2 - x1 * x1));
        y2 = sqrt(fabs(y2 - y1 * y1));
        double farfield = v[5] * v[5] + v[6] * v[6];
        farfield *= dgrid * dgrid;
        double inten = v[7];
        g_ff += farfield;
        g_in += inten;
        power *= scl * scl / kVacimp;
        inten *= kEev * kEev / ks / ks / kVacimp;
        if (o->power) o->power[is] = power;
        if (o->f_xposition) o->f_xposition[is] = x1 * dgrid;
        if (o->f_xsize) o->f_xsize[is] = x2 * dgrid;
        if (o->f_yposition) o->f_yposition[is] = y1 * dgrid;
        if (o->f_ysize) o->f_ysize[is] = y2 * dgrid;
        if (o->nf_intensity) o->nf_intensity[is] = inten;
        if (o->ff_intensity) o->ff_intensity[is] = farfield;
    }
    {
        const double norm = g_pow > 0 ? 1. / g_pow : 1.;
        fx1 *= norm, fx2 *= norm, fy1 *= norm, fy2 *= norm;
        if (o->g_f_xsize) o->g_f_xsize[0] = sqrt(fabs(fx2 - fx1 * fx1)) * dgrid;
        if (o->g_f_ysize) o->g_f_ysize[0] = sqrt(fabs(fy2 - fy1 * fy1)) * dgrid;
        if (o->g_f_xposition) o->g_f_xposition[0] = fx1 * dgrid;
        if (o->g_f_yposition) o->g_f_yposition[0] = fy1 * dgrid;
        const double ninv = 1. / (double)ns;
        if (o->g_nf_intensity) o->g_nf_intensity[0] = g_in * ninv * kEev * kEev / ks / ks / kVacimp;
        if (o->g_ff_intensity) o->g_ff_intensity[0] = g_ff * ninv;
    }
    if (o->global_energy) {
        g_pow *= scl * scl / kVacimp * slicelength / 299792458.0;
        o->global_energy[0] = g_pow;
    }
}

static void finish_diag(const Ctx *c, const std::vector<double> &hb, const std::vector<double> &hf,
                        int first, g4mi_diag_out *o) {
    const int ns = c->ns;
    // DiagBeam::getValues normalisation (Diagnostic.cpp:452-489) and the
    // current-weighted sums of the global statistics (:552-607)
    double gc = 0, gg1 = 0, gg2 = 0, gx1 = 0, gx2 = 0, gy1 = 0, gy2 = 0;
    for (int is = 0; is < ns; is++) {
        const long long n = c->off[is + 1] - c->off[is];
        const double norm = n > 0 ? 1. / (double)n : 1.;
        const double *v = &hb[8 * is];
        const double g1 = v[0] * norm, g2 = v[1] * norm, x1 = v[2] * norm, x2 = v[3] * norm;
        const double y1 = v[4] * norm, y2 = v[5] * norm, bre = v[6] * norm, bim = v[7] * norm;
        const double cu = is < (int)c->h_cur.size() ? c->h_cur[is] : 0.0;
        gc += cu;
        gg1 += cu * g1;
        gg2 += cu * g2;
        gx1 += cu * x1;
        gx2 += cu * x2;
        gy1 += cu * y1;
        gy2 += cu * y2;
        if (o->energy) o->energy[is] = g1;
        if (o->energyspread) o->energyspread[is] = sqrt(fabs(g2 - g1 * g1));
        if (o->xposition) o->xposition[is] = x1;
        if (o->xsize) o->xsize[is] = sqrt(fabs(x2 - x1 * x1));
        if (o->yposition) o->yposition[is] = y1;
        if (o->ysize) o->ysize[is] = sqrt(fabs(y2 - y1 * y1));
        if (o->bunching) o->bunching[is] = hypot(bre, bim);
        if (o->bunchingphase) o->bunchingphase[is] = atan2(bim, bre);
    }
    {
        const double norm = gc > 0 ? 1. / gc : 1.;
        gg1 *= norm, gg2 *= norm, gx1 *= norm, gx2 *= norm, gy1 *= norm, gy2 *= norm;
        if (o->g_energy) o->g_energy[0] = gg1;
        if (o->g_energyspread) o->g_energyspread[0] = sqrt(fabs(gg2 - gg1 * gg1));
        if (o->g_xposition) o->g_xposition[0] = gx1;
        if (o->g_xsize) o->g_xsize[0] = sqrt(fabs(gx2 - gx1 * gx1));
        if (o->g_yposition) o->g_yposition[0] = gy1;
        if (o->g_ysize) o->g_ysize[0] = sqrt(fabs(gy2 - gy1 * gy1));
    }
    finish_field(hf, ns, first, c->dgrid, 4. * asin(1) / c->xlambda, c->slicelength, o);
}

// pinned record slots (at most 32 MiB, 4..256 records), allocated with the
// context so no run pays for the allocation
// A deferred record's slot: beam and field sums (16 per slice), the
// auxiliary sums (16), the long- and short-range space-charge fields (1 + 1)
// and the spectral moments (5).
constexpr int kRecRow = 39;

static int alloc_records(Ctx *c) {
    const size_t row = kRecRow * (size_t)c->ns;
    c->rec_cap = (int)std::max<size_t>(4, std::min<size_t>(256, ((size_t)32 << 20) / (row * sizeof(double))));
    G4_HIP(hipHostMalloc((void **)&c->h_recs, (size_t)c->rec_cap * row * sizeof(double)));
    G4_HIP(hipHostGetDevicePointer((void **)&c->d_recs, c->h_recs, 0));
    return G4MI_OK;
}

static bool wants_aux(const g4mi_diag_out *o) {
    return o->pxposition || o->pyposition || o->nf_phase || o->efield || o->wakefield || o->lscfield ||
           o->sscfield || o->xmin || o->xmax || o->pxmin || o->pxmax || o->ymin || o->ymax || o->pymin ||
           o->pymax || o->emin || o->emax;
}

// The auxiliary outputs (DiagBeam :415-445, :505-519; DiagField :855-858)
// from aux[16 ns] of launch_diag_aux / launch_mix_diag_aux and the per-slice
// long-range (esc) and short-range (ssc) space-charge fields.
static void finish_aux(const Ctx *c, const double *aux, const double *esc, const double *ssc,
                       const std::vector<double> &wake, g4mi_diag_out *o) {
    double *dst[10] = {o->xmin, o->xmax, o->pxmin, o->pxmax, o->ymin, o->ymax, o->pymin, o->pymax, o->emin, o->emax};
    for (int is = 0; is < c->ns; is++) {
        const long long n = c->off[is + 1] - c->off[is];
        const double norm = n > 0 ? 1. / (double)n : 1.;
        const double *a = &aux[16 * (size_t)is];
        if (o->pxposition) o->pxposition[is] = a[0] * norm;
        if (o->pyposition) o->pyposition[is] = a[1] * norm;
        for (int k = 0; k < 10; k++)
            if (dst[k]) dst[k][is] = a[2 + k];
        if (o->nf_phase) o->nf_phase[is] = atan2(a[13], a[12]);
        // Beam::eloss: the wake loss of the last step that applied it
        const double eloss = is < (int)wake.size() ? wake[is] : 0.0;
        if (o->efield) o->efield[is] = eloss + esc[is];
        if (o->wakefield) o->wakefield[is] = eloss;
        if (o->lscfield) o->lscfield[is] = esc[is];
        if (o->sscfield) o->sscfield[is] = ssc[is];
    }
}

static MixArrays mix_arrays(const Ctx *c);

// the aux sums of this step into d_aux (device)
static int launch_aux(Ctx *c) {
    const int ns = c->ns;
    if (!c->d_aux) G4_TRY(dalloc(c, &c->d_aux, 16 * (size_t)ns));
    if (c->precision == G4MI_FP32)
        launch_mix_diag_aux(c->stream, mix_arrays(c), c->d_aux);
    else
        launch_diag_aux(c->stream, ns, c->d_off, c->d_g, c->d_x, c->d_y, c->d_px, c->d_py, c->d_field, c->ng,
                        c->first, c->d_aux);
    return G4MI_OK;
}

// aux sums plus the space-charge fields, copied to the host
static int fetch_aux(Ctx *c, std::vector<double> &aux, std::vector<double> &esc, std::vector<double> &ssc) {
    const int ns = c->ns;
    G4_TRY(launch_aux(c));
    aux.resize(16 * (size_t)ns);
    esc.resize(ns);
    ssc.resize(ns);
    G4_HIP(hipMemcpyAsync(aux.data(), c->d_aux, aux.size() * 8, hipMemcpyDeviceToHost, c->stream));
    G4_HIP(hipMemcpyAsync(esc.data(), c->d_esc, esc.size() * 8, hipMemcpyDeviceToHost, c->stream));
    G4_HIP(hipMemcpyAsync(ssc.data(), c->d_efield, ssc.size() * 8, hipMemcpyDeviceToHost, c->stream));
    G4_HIP(hipStreamSynchronize(c->stream));
    return G4MI_OK;
}

// DFT matrices W, W* and two field-sized work buffers of the spectral
// solver's filter and the spectral diagnostics, allocated on first use.
static int ensure_dft(Ctx *c) {
    if (c->d_fft_w) return G4MI_OK;
    const int ng = c->ng;
    const size_t nn = (size_t)ng * ng, cells = (size_t)c->ns * nn;
    G4_TRY(dalloc(c, &c->d_fft_w, nn));
    G4_TRY(dalloc(c, &c->d_fft_wc, nn));
    G4_TRY(dalloc(c, &c->d_fft_u, cells));
    G4_TRY(dalloc(c, &c->d_fft_v, cells));
    G4_TRY(dalloc(c, &c->d_fft_mom, 5 * (size_t)c->ns));
    G4_TRY(dalloc(c, &c->d_fft_mompart, 5 * (size_t)c->ns * fft_moments_parts(c->ns)));
    std::vector<double> w(2 * nn), wc(2 * nn);
    dft_matrix(ng, -1, w.data());
    dft_matrix(ng, +1, wc.data());
    G4_HIP(hipMemcpy(c->d_fft_w, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice));
    G4_HIP(hipMemcpy(c->d_fft_wc, wc.data(), wc.size() * sizeof(double), hipMemcpyHostToDevice));
    return G4MI_OK;
}

static bool wants_spectral(const g4mi_diag_out *o) {
    return o->xdivergence || o->ydivergence || o->xpointing || o->ypointing || o->g_xdivergence ||
           o->g_ydivergence || o->g_xpointing || o->g_ypointing;
}

// The twiddles and the Bluestein tables of the DFTs by FFT convolution
// (ngrid <= 256), computed once per context.
static int ensure_bluestein(Ctx *c) {
    const int L = fft_conv_length(c->ng);
    if (!L) return G4MI_E_ARG;
    if (c->d_fft_bf) return G4MI_OK;
    if (!c->d_fft_tw) {
        G4_TRY(dalloc(c, &c->d_fft_hk, 512));
        G4_TRY(dalloc(c, &c->d_fft_tw, 512));
        std::vector<double> hk(2 * (size_t)L), tw(2 * (size_t)L);
        (void)fft_conv_tables(c->ng, c->dgrid, c->xks, 1.0, hk.data(), tw.data());  // twiddles only
        G4_HIP(hipMemcpy(c->d_fft_tw, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    double2 *bf = nullptr, *bi = nullptr;
    G4_TRY(dalloc(c, &bf, 512 + 256));
    G4_TRY(dalloc(c, &bi, 512 + 256));
    std::vector<double> hk(2 * (size_t)L), ch(2 * (size_t)c->ng);
    for (int sign : {-1, +1}) {
        (void)fft_chirp_tables(c->ng, sign, hk.data(), ch.data());
        double2 *dst = sign < 0 ? bf : bi;
        G4_HIP(hipMemcpy(dst, hk.data(), hk.size() * sizeof(double), hipMemcpyHostToDevice));
        G4_HIP(hipMemcpy(dst + 512, ch.data(), ch.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    c->d_fft_bf = bf;
    c->d_fft_bi = bi;
    return G4MI_OK;
}

// One DFT pass (rows or columns) of every slice by Bluestein: sign -1
// forward, +1 inverse (unnormalised).
static ConvArgs bluestein_pass(const Ctx *c, int sign, const double2 *in, double2 *out) {
    ConvArgs a{};
    const double2 *t = sign < 0 ? c->d_fft_bf : c->d_fft_bi;
    a.in = in;
    a.out = out;
    a.hk = t;
    a.pre = a.post = t + 512;
    a.tw = c->d_fft_tw;
    a.nrows = (long long)c->ns * c->ng;
    a.ng = c->ng;
    a.ns = c->ns;
    a.L = fft_conv_length(c->ng);
    return a;
}

// The moments of |DFT(E)|^2 per slice into d_fft_mom (device, 5 ns).
static int spectral_moments(Ctx *c, const double2 *wide) {
    G4_TRY(ensure_dft(c));
    const int ns = c->ns, ng = c->ng;
    // the products cannot run in place: the intermediate is the work buffer
    // that does not hold the input, the result goes to the other one
    double2 *inter = wide == c->d_fft_u ? c->d_fft_v : c->d_fft_u;
    double2 *res = wide == c->d_fft_u ? c->d_fft_u : c->d_fft_v;
    if (fft_conv_length(ng) && c->fft_prop != 0) {  // DFTs by Bluestein unless the GEMMs are asked for
        G4_TRY(ensure_bluestein(c));
        G4_TRY(launch_fft_conv_rows(c->stream, bluestein_pass(c, -1, wide, inter)));
        G4_TRY(launch_fft_conv_cols(c->stream, bluestein_pass(c, -1, inter, res)));
    } else {
        CgemmArgs g{};
        g.ng = ng;
        g.ns = ns;
        g.nrows = (long long)ns * ng;
        g.scale = 1.0;
        g.a = wide;
        g.mat = c->d_fft_w;
        g.c = inter;
        G4_TRY(launch_cgemm_rows(c->stream, g));
        g.b = inter;
        g.c = res;
        G4_TRY(launch_cgemm_cols(c->stream, g));
    }
    launch_fft_moments(c->stream, ns, res, ng, c->d_fft_mompart, c->d_fft_mom);
    return G4MI_OK;
}

// DiagField's spectral outputs from the moments m[5 ns] (storage order),
// normalised on the host.
static void finish_spectral(const Ctx *c, const double *m, int first, g4mi_diag_out *o) {
    const int ns = c->ns, ng = c->ng;
    const double scltheta = c->xlambda / ng / c->dgrid;
    double f[5] = {0, 0, 0, 0, 0};
    for (int k = 0; k < ns; k++) {  // storage slice k is output slice (ns + k - first) mod ns
        const int is = (ns + k - first) % ns;
        double v[5];
        for (int q = 0; q < 5; q++) {
            v[q] = m[5 * (size_t)k + q];
            f[q] += v[q];
        }
        if (v[0] > 0)
            for (int q = 1; q < 5; q++) v[q] /= v[0];
        if (o->xdivergence) o->xdivergence[is] = sqrt(fabs(v[2] - v[1] * v[1])) * scltheta;
        if (o->ydivergence) o->ydivergence[is] = sqrt(fabs(v[4] - v[3] * v[3])) * scltheta;
        if (o->xpointing) o->xpointing[is] = v[1] * scltheta;
        if (o->ypointing) o->ypointing[is] = v[3] * scltheta;
    }
    const double norm = f[0] > 0 ? 1. / f[0] : 1.;
    for (int q = 1; q < 5; q++) f[q] *= norm;
    if (o->g_xdivergence) o->g_xdivergence[0] = sqrt(fabs(f[2] - f[1] * f[1])) * scltheta;
    if (o->g_ydivergence) o->g_ydivergence[0] = sqrt(fabs(f[4] - f[3] * f[3])) * scltheta;
    if (o->g_xpointing) o->g_xpointing[0] = f[1] * scltheta;
    if (o->g_ypointing) o->g_ypointing[0] = f[3] * scltheta;
}

// DiagField's spectral block (src/Core/Diagnostic.cpp:789-807, 839-846 and
// the global values :952-963): F = W E W^T per slice, moments of |F|^2 per
// slice, normalised on the host.  `wide` is the FP64 field ([ns][ng][ng],
// storage order), `first` the ring offset.
static int spectral_diag(Ctx *c, const double2 *wide, int first, g4mi_diag_out *o) {
    G4_TRY(spectral_moments(c, wide));
    std::vector<double> m(5 * (size_t)c->ns);
    G4_HIP(hipMemcpyAsync(m.data(), c->d_fft_mom, m.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    G4_HIP(hipStreamSynchronize(c->stream));
    finish_spectral(c, m.data(), first, o);
    return G4MI_OK;
}

static int diag(Ctx *c, g4mi_diag_out *o) {
    const int ns = c->ns, ng = c->ng;
    launch_diag_beam(c->stream, ns, c->d_off, c->d_g, c->d_th, c->d_x, c->d_y, c->d_diag_b);
    // the fused ADI already summed the field of this step; otherwise sum now
    if (!c->field_diag_fresh) launch_diag_field(c->stream, ns, c->d_field, ng, c->d_diag_f);
    std::vector<double> hb(8 * (size_t)ns), hf(8 * (size_t)ns);
    G4_HIP(hipMemcpyAsync(hb.data(), c->d_diag_b, hb.size() * 8, hipMemcpyDeviceToHost, c->stream));
    G4_HIP(hipMemcpyAsync(hf.data(), c->d_diag_f, hf.size() * 8, hipMemcpyDeviceToHost, c->stream));
    G4_HIP(hipStreamSynchronize(c->stream));
    if (c->field_diag_fresh)  // slices zeroed by slippage after the ADI sum to 0
        for (int z : c->zeroed_since_adi)
            for (int k = 0; k < 8; k++) hf[8 * (size_t)z + k] = 0;
    finish_diag(c, hb, hf, c->first, o);
    if (wants_aux(o)) {
        std::vector<double> aux, esc, ssc;
        G4_TRY(fetch_aux(c, aux, esc, ssc));
        finish_aux(c, aux.data(), esc.data(), ssc.data(), c->wake_applied, o);
    }
    if (wants_spectral(o)) G4_TRY(spectral_diag(c, c->d_field, c->first, o));
    return G4MI_OK;
}

// Fold the recorded events of the pool into the per-phase totals.
static int fold_timers(Ctx *c) {
    if (c->pool_used == 0) return G4MI_OK;
    G4_HIP(hipStreamSynchronize(c->stream));
    for (int k = 0; k < c->pool_used; k++) {
        hipEvent_t *e = &c->pool[(size_t)k * Ctx::kEv];
        float d[8] = {};
        for (int j = 0; j < 8; j++) G4_HIP(hipEventElapsedTime(&d[j], e[j], e[j + 1]));
        // e0 trans1 e1 sc e2 push e3 trans2 e4 host e5 deposit e6 adi e7 slip+diag e8
        c->t_phase[0] += d[0] + d[3];
        c->t_phase[1] += d[1];
        c->t_phase[2] += d[2];
        c->t_phase[3] += d[5];
        c->t_phase[4] += d[6];
        c->t_phase[5] += d[7];
        c->n_timed++;
    }
    c->pool_used = 0;
    return G4MI_OK;
}

static int step_begin_timing(Ctx *c) {
    c->ev = nullptr;
    if (!c->timing) return G4MI_OK;
    if (c->pool_used == Ctx::kPoolSteps) G4_TRY(fold_timers(c));
    c->ev = &c->pool[(size_t)c->pool_used * Ctx::kEv];
    return G4MI_OK;
}

static int step_end_timing(Ctx *c) {
    if (!c->ev) return G4MI_OK;
    G4_HIP(hipEventRecord(c->ev[8], c->stream));
    c->ev = nullptr;
    c->pool_used++;
    return G4MI_OK;
}

// Beam::sort (src/Core/Beam.cpp:81-94) of a one4one beam: migration on the
// device, then the host mirror follows the new slice offsets: currents
// np ce / slicelength, the push block table and the incoherent beamlets.
static int sort_beam(Ctx *c) {
    const size_t n = (size_t)std::max(1LL, c->ntot);
    if (!c->sw.keys) {
        SortWork &w = c->sw;
        G4_TRY(dalloc(c, &w.keys, n));
        G4_TRY(dalloc(c, &w.keys_out, n));
        G4_TRY(dalloc(c, &w.vals, n));
        G4_TRY(dalloc(c, &w.vals_out, n));
        G4_TRY(dalloc(c, &w.th_new, n));
        for (double **p : {&w.g, &w.th, &w.x, &w.y, &w.px, &w.py}) G4_TRY(dalloc(c, p, n));
        G4_TRY(sort_work_bytes(c->ns, c->ntot, &w.temp_bytes));
        G4_TRY(dalloc(c, (char **)&w.temp, std::max<size_t>(1, w.temp_bytes)));
        G4_TRY(dalloc(c, &c->d_off_new, (size_t)c->ns + 1));
    }
    const double slen = 4. * asin(1.) * c->slicelength / c->lambda0;  // Beam::initSorting (:75-79)
    G4_TRY(launch_sort_one4one(c->stream, c->ns, c->ntot, c->d_off, slen, c->d_g, c->d_th, c->d_x, c->d_y,
                               c->d_px, c->d_py, c->sw, c->d_off_new));
    std::swap(c->d_g, c->sw.g);
    std::swap(c->d_th, c->sw.th);
    std::swap(c->d_x, c->sw.x);
    std::swap(c->d_y, c->sw.y);
    std::swap(c->d_px, c->sw.px);
    std::swap(c->d_py, c->sw.py);
    std::swap(c->d_off, c->d_off_new);
    G4_HIP(hipMemcpyAsync(c->off.data(), c->d_off, (c->ns + 1) * sizeof(long long), hipMemcpyDeviceToHost,
                          c->stream));
    G4_HIP(hipStreamSynchronize(c->stream));
    c->ntot = c->off[c->ns];
    c->h_cur.assign(c->ns, 0.0);
    for (int s = 0; s < c->ns; s++) c->h_cur[s] = (double)(c->off[s + 1] - c->off[s]) * kCe / c->slicelength;
    std::vector<int> bs;
    std::vector<long long> bb;
    build_block_table(c->off, bs, bb);
    if ((int)bs.size() > c->blk_cap) {
        set_error("sort: block table capacity exceeded");
        return G4MI_E_STATE;
    }
    c->nblk = (int)bs.size();
    G4_HIP(hipMemcpyAsync(c->d_cur, c->h_cur.data(), c->ns * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (c->nblk) {
        G4_HIP(hipMemcpyAsync(c->d_blk_slice, bs.data(), bs.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
        G4_HIP(hipMemcpyAsync(c->d_blk_begin, bb.data(), bb.size() * sizeof(long long), hipMemcpyHostToDevice,
                              c->stream));
    }
    if (c->incoh) {
        std::vector<long long> boff(c->ns + 1, 0);
        for (int s = 0; s < c->ns; s++)
            boff[s + 1] = boff[s] + (c->off[s + 1] - c->off[s] + c->incoh_nbins - 1) / c->incoh_nbins;
        c->incoh_nb = boff[c->ns];
        G4_HIP(hipMemcpyAsync(c->d_boff, boff.data(), boff.size() * sizeof(long long), hipMemcpyHostToDevice,
                              c->stream));
    }
    G4_HIP(hipStreamSynchronize(c->stream));  // the host vectors above go out of scope
    return G4MI_OK;
}

// Whether a step can run as the two fused particle kernels: fundamental
// field only, no energy kicks (incoherent SR, wakes) between the push and
// R56, no sort between the beam and the field track, and every space-charge
// mode solution within the main kernel's LDS.
constexpr int kFuseMinSlices = 256;  // one workgroup per CU at least
static bool fp64_fuse_on(const Ctx *c) { return c->fuse64 > 0 || (c->fuse64 < 0 && c->ns >= kFuseMinSlices); }
static bool fp64_fusable(const Ctx *c, const g4mi_step_params &st) {
    if (!fp64_fuse_on(c) || !c->hf.empty() || c->d_wake) return false;
    if (c->incoh && st.aw > 0 && (c->incoh_loss || c->incoh_spread)) return false;
    if ((st.marker & 4) && c->one4one) return false;
    const bool sc = (c->sc_nz > 0) && (c->sc_ng > 2);
    if (sc && (c->sc_ng > 256 || (2 * c->sc_nphi + 1) * c->sc_nz * c->sc_ng > fp64_main_modes())) return false;
    return true;
}

// Beam::track + Field::track of one step with the fused kernels: front
// (first half step, moments, radius, footprint), the long-range field and
// the radial scan, main (short range, push, R56, second half step, beam
// sums, deposit), then the field solve.
static int fused_step(Ctx *c, const g4mi_step_params &st, double gamma0, double delz) {
    if (!c->has_beam) {
        set_error("beam not uploaded");
        return G4MI_E_STATE;
    }
    const double aw = st.aw;
    const bool sc = (c->sc_nz > 0) && (c->sc_ng > 2);
    const bool moments = c->longrange || sc;
    TransParams t1 = transverse_params(c, st, gamma0, delz * 0.5, false);
    FuseParams f{};
    PushParams &p = f.p;
    p.delz = delz;
    p.xks = c->xks;
    p.rtmp = coupling(st, 1) / c->xks;
    p.xku = st.ku;
    if (p.xku == 0) p.xku = c->xks * 0.5 / gamma0 / gamma0;
    p.aw = aw;
    p.autophase = st.phaseshift;
    p.ax = st.ax;
    p.ay = st.ay;
    p.kx = st.kx;
    p.ky = st.ky;
    p.gradx = st.gradx;
    p.grady = st.grady;
    p.gridmax = c->gridmax;
    p.dgrid = c->dgrid;
    p.ngrid = c->ng;
    p.first = c->first;
    p.nslice = c->ns;
    p.use_sc = sc ? 1 : 0;
    p.nh = 0;
    f.t2 = transverse_params(c, st, gamma0, delz * 0.5, true);
    DepositParams &d = f.d;
    d.enabled = st.aw > 0 ? 1 : 0;  // in undulator, fundamental
    d.harm = 1.0;
    d.fcoup = coupling(st, 1);
    d.xks = c->xks;
    d.delz = delz;
    d.dgrid = c->dgrid;
    d.gridmax = c->gridmax;
    d.ax = st.ax;
    d.ay = st.ay;
    d.kx = st.kx;
    d.ky = st.ky;
    d.gradx = st.gradx;
    d.grady = st.grady;
    d.ngrid = c->ng;
    f.chicane = st.chic_angle != 0 ? 1 : 0;
    if (f.chicane) f.r56 = chicane_r56(st.chic_angle, st.chic_lb, st.chic_ld, c->lambda0, gamma0);
    f.gamma0 = gamma0;
    f.use_sc = sc ? 1 : 0;
    f.ngr = c->sc_ng;
    f.nz = c->sc_nz;
    f.nphi = c->sc_nphi;
    f.ks = 4 * asin(1) / c->lambda0;
    f.gz2 = gamma0 * gamma0 / (1 + aw * aw);
    if (c->ev) G4_HIP(hipEventRecord(c->ev[0], c->stream));
    launch_fp64_front(c->stream, c->ns, c->d_off, c->d_g, c->d_x, c->d_y, c->d_px, c->d_py, c->d_cur, t1, moments,
                      sc, c->gridmax, c->dgrid, c->ng, c->d_mom, c->d_fsize, c->d_cen, c->d_rbound, c->d_pbox,
                      f.d, c->d_pmaxf);
    if (c->ev) G4_HIP(hipEventRecord(c->ev[1], c->stream));
    if (c->longrange) {
        const double gamma = gamma0 / sqrt(1 + aw * aw);
        const double scl = c->slicelength / 2. / asin(1) / 299792458.0 / 2 / 8.85e-12;
        launch_lsc_long(c->stream, c->ns, c->d_cur, c->d_fsize, c->slicelength, gamma, scl, c->d_esc);
    }
    if (sc) launch_sc_scan(c->stream, c->ns, c->d_off, c->d_rbound, c->d_rmax, c->sc_ng, c->d_dr);
    if (c->ev) G4_HIP(hipEventRecord(c->ev[2], c->stream));
    launch_fp64_main(c->stream, c->ns, c->d_off, c->d_g, c->d_th, c->d_x, c->d_y, c->d_px, c->d_py, c->d_field,
                     c->d_esc, c->d_cur, c->d_cen, c->d_dr, c->d_efield, c->d_pbox, c->d_pmaxf, c->d_src,
                     c->d_box, c->d_beamsum, f);
    if (c->ev) {
        G4_HIP(hipEventRecord(c->ev[3], c->stream));
        G4_HIP(hipEventRecord(c->ev[4], c->stream));
    }
    G4_HIP(hipGetLastError());
    G4_TRY(field_track(c, st, delz, false));
    c->beamsum_fresh = true;
    return G4MI_OK;
}

static int step(Ctx *c, const g4mi_step_params &st, double gamma0) {
    const double delz = st.dz;
    c->beamsum_fresh = false;
    G4_TRY(step_begin_timing(c));
    if (fp64_fusable(c, st)) {
        G4_TRY(fused_step(c, st, gamma0, delz));
    } else {
        G4_TRY(beam_track(c, st, gamma0, delz));
        if ((st.marker & 4) && c->one4one) G4_TRY(sort_beam(c));  // Gencore.cpp:124-134
        G4_TRY(field_track(c, st, delz));
    }
    G4_TRY(slippage(c, st.slip));
    return G4MI_OK;
}

// ===========================================================================
// Mixed form (precision G4MI_FP32): device step table, banded FP32 inverse
// per step size, one graph per launch signature replayed for every step.
// ===========================================================================
constexpr int kMixMaxBands = 64;
constexpr int kMixBandRowMax = 2 * 12 + 1;
constexpr int kMixRecRows = 32;

static MixArrays mix_arrays(const Ctx *c) {
    MixArrays a{};
    a.dg = c->m_dg;
    a.dth = c->m_dth;
    a.x = c->m_x;
    a.y = c->m_y;
    a.px = c->m_px;
    a.py = c->m_py;
    a.xs = c->m_xs;
    a.ys = c->m_ys;
    a.base = c->m_base;
    a.off = c->d_off;
    a.cur = c->d_cur;
    a.field[0] = c->m_field[0];
    a.field[1] = c->m_field[1];
    a.src = c->m_src;
    a.box = c->d_box;
    a.pbox = c->m_pbox;
    a.mom_fsize[0] = c->d_fsize;
    a.mom_fsize[1] = c->m_mom_next;
    a.cen = c->d_cen;
    a.rbound = c->d_rbound;
    a.dr = c->d_dr;
    a.rmax = c->d_rmax;
    a.esc = c->d_esc;
    a.efield = c->d_efield;
    a.beamsum = c->m_beamsum;
    a.egsum = c->m_egsum;
    a.tilediag = c->m_tilediag;
    a.rec = c->m_rec;
    a.nrec = c->m_nrec;
    a.steps = c->m_steps;
    a.st = c->m_st;
    a.bands = c->m_bands;
    a.binfo = c->m_binfo;
    a.ns = c->ns;
    a.ng = c->ng;
    a.sc_ng = c->sc_ng;
    a.sc_nz = c->sc_nz;
    a.sc_nphi = c->sc_nphi;
    a.gref = c->gref;
    a.gridmax = c->gridmax;
    a.dgrid = c->dgrid;
    a.sample = c->sample;
    a.timerun = c->timerun;
    a.periodic = c->periodic;
    return a;
}

static void mix_clear_graphs(Ctx *c) {
    for (auto &kv : c->m_graph) (void)hipGraphExecDestroy(kv.second);
    c->m_graph.clear();
}

// FP32 banded inverse for step size delz: the narrowest half width whose
// dropped entries are below 2^-28 of the diagonal, stored with the width of
// the kernel instantiation that runs it; cached per step size.
static int mix_band(Ctx *c, double delz, int *band) {
    auto it = c->band_of_dz.find(delz);
    if (it != c->band_of_dz.end()) {
        *band = it->second;
        return G4MI_OK;
    }
    if ((int)c->h_binfo.size() >= kMixMaxBands) {
        set_error("mixed form: more distinct step sizes than the band table holds (64)");
        return G4MI_E_ARG;
    }
    const int ng = c->ng;
    std::vector<double> h(6 * (size_t)ng);
    const double rtmp = adi_factors(ng, delz, c->dgrid, c->xks, h.data(), h.data() + 2 * ng,
                                    h.data() + 4 * ng);
    int need = -1, tm = 0;
    for (int m = 1; m <= 12 && need < 0; m++) {
        std::vector<double> tmp(2 * (size_t)ng * (2 * m + 1));
        if (adi_band(ng, h.data(), h.data() + 2 * ng, h.data() + 4 * ng, m, tmp.data()) <
            std::ldexp(1.0, -28))
            need = m;
    }
    if (need < 0 || !mix_band_supported(need, &tm)) {
        set_error("mixed form: the step size needs a band wider than 12 cells");
        return G4MI_E_ARG;
    }
    const int W = 2 * tm + 1;
    std::vector<double> bd(2 * (size_t)ng * W);
    (void)adi_band(ng, h.data(), h.data() + 2 * ng, h.data() + 4 * ng, tm, bd.data());
    // interior rows whose band equals the central row's to 2^-30 (FP32)
    const int kc = ng / 2;
    auto same = [&](int k) {
        const double dia = std::hypot(bd[2 * ((size_t)kc * W + tm)], bd[2 * ((size_t)kc * W + tm) + 1]);
        double d = 0;
        for (int q = 0; q < W; q++)
            d = std::max(d, std::hypot(bd[2 * ((size_t)k * W + q)] - bd[2 * ((size_t)kc * W + q)],
                                       bd[2 * ((size_t)k * W + q) + 1] - bd[2 * ((size_t)kc * W + q) + 1]));
        return d <= std::ldexp(1.0, -30) * dia;
    };
    int lo = kc, hi = kc;
    while (lo > 0 && same(lo - 1)) lo--;
    while (hi < ng - 1 && same(hi + 1)) hi++;
    const int half = std::min(kc - lo, hi - kc);
    MixBand bi{};
    bi.off = (long long)c->h_bands.size();
    bi.a = rtmp;
    bi.klo = kc - half;
    bi.khi = kc + half;
    bi.m = need;
    for (size_t k = 0; k < bd.size() / 2; k++)
        c->h_bands.push_back(make_float2((float)bd[2 * k], (float)bd[2 * k + 1]));
    const int idx = (int)c->h_binfo.size();
    c->h_binfo.push_back(bi);
    c->band_tmpl.push_back(tm);
    G4_HIP(hipMemcpyAsync(c->m_bands + bi.off, c->h_bands.data() + bi.off, bd.size() / 2 * sizeof(float2),
                          hipMemcpyHostToDevice, c->stream));
    G4_HIP(hipMemcpyAsync(c->m_binfo + idx, &c->h_binfo[idx], sizeof(MixBand), hipMemcpyHostToDevice,
                          c->stream));
    G4_HIP(hipStreamSynchronize(c->stream));
    c->band_of_dz[delz] = idx;
    *band = idx;
    return G4MI_OK;
}

static int mix_fill_step(Ctx *c, const g4mi_step_params &st, double gamma0, MixStep &m) {
    const double delz = st.dz;
    m = MixStep{};
    m.t1 = transverse_params(c, st, gamma0, delz * 0.5, false);
    m.t2 = transverse_params(c, st, gamma0, delz * 0.5, true);
    PushParams &p = m.p;
    p.delz = delz;
    p.xks = c->xks;
    p.rtmp = coupling(st, 1) / c->xks;
    p.xku = st.ku;
    if (p.xku == 0) p.xku = c->xks * 0.5 / gamma0 / gamma0;
    p.aw = st.aw;
    p.autophase = st.phaseshift;
    p.ax = st.ax;
    p.ay = st.ay;
    p.kx = st.kx;
    p.ky = st.ky;
    p.gradx = st.gradx;
    p.grady = st.grady;
    p.gridmax = c->gridmax;
    p.dgrid = c->dgrid;
    p.ngrid = c->ng;
    DepositParams &d = m.d;
    d.enabled = st.aw > 0 ? 1 : 0;
    d.fcoup = coupling(st, 1);
    d.xks = c->xks;
    d.delz = delz;
    d.dgrid = c->dgrid;
    d.gridmax = c->gridmax;
    d.ax = st.ax;
    d.ay = st.ay;
    d.kx = st.kx;
    d.ky = st.ky;
    d.gradx = st.gradx;
    d.grady = st.grady;
    d.ngrid = c->ng;
    m.gamma0 = gamma0;
    m.chicane = st.chic_angle != 0 ? 1 : 0;
    if (m.chicane) m.r56 = chicane_r56(st.chic_angle, st.chic_lb, st.chic_ld, c->lambda0, gamma0);
    // reference particle: gamma = gref, on axis, no transverse momentum, no field
    m.eps_ref = (1 + st.aw * st.aw) / (c->gref * c->gref);
    // xks (1 - 1/beta) + xku with 1 - 1/beta = -eps / (beta (1 + beta)), no cancellation
    const double bref = sqrt(1. - m.eps_ref);
    m.kappa = -c->xks * m.eps_ref / (bref * (1. + bref)) + p.xku;
    m.longrange = c->longrange;
    if (c->longrange) {
        const double gamma = gamma0 / sqrt(1 + st.aw * st.aw);
        m.lsc_ds = c->slicelength * gamma;
        m.lsc_scl = c->slicelength / 2. / asin(1) / 299792458.0 / 2 / 8.85e-12;
    }
    m.use_sc = (c->sc_nz > 0 && c->sc_ng > 2) ? 1 : 0;
    m.sc_gz2 = gamma0 * gamma0 / (1 + st.aw * st.aw);
    m.sc_ks = 4 * asin(1) / c->lambda0;
    m.slip = st.slip;
    G4_TRY(mix_band(c, delz, &m.band));
    return G4MI_OK;
}

// band, refine, sc, moments, maxzero, ADI kernel (kAdi*), main threads per slice, main LDS region,
// front kernel launched (0 when the previous step's main kernel ran the front)
typedef std::tuple<int, int, int, int, int, int, int, int, int> MixKey;

static int mix_launch(Ctx *c, const MixKey &k, hipEvent_t *ev) {
    const MixArrays a = mix_arrays(c);
    if (launch_mix_step(c->stream, a, std::get<0>(k), std::get<1>(k) != 0, std::get<2>(k) != 0,
                        std::get<3>(k) != 0, std::get<4>(k), ev, std::get<5>(k), std::get<6>(k),
                        std::get<7>(k), std::get<8>(k) != 0) != 0) {
        set_error("mixed form: no ADI instantiation for this band width");
        return G4MI_E_ARG;
    }
    G4_HIP(hipGetLastError());
    return G4MI_OK;
}

static int mix_graph(Ctx *c, const MixKey &k, hipGraphExec_t *out) {
    auto it = c->m_graph.find(k);
    if (it != c->m_graph.end()) {
        *out = it->second;
        return G4MI_OK;
    }
    hipGraph_t g = nullptr;
    G4_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int rc = mix_launch(c, k, nullptr);
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc != G4MI_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    G4_HIP(e);
    hipGraphExec_t ex = nullptr;
    e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    G4_HIP(e);
    c->m_graph[k] = ex;
    *out = ex;
    return G4MI_OK;
}

// DiagBeam/DiagField normalisation of one record row (sums per beam slice).
static void mix_finish_row(const Ctx *c, const double *row, g4mi_diag_out *o) {
    const int ns = c->ns;
    std::vector<double> hb(8 * (size_t)ns), hf(8 * (size_t)ns);
    for (int is = 0; is < ns; is++)
        for (int q = 0; q < 8; q++) {
            hb[8 * (size_t)is + q] = row[(size_t)is * kMixRecord + q];
            hf[8 * (size_t)is + q] = row[(size_t)is * kMixRecord + 8 + q];
        }
    finish_diag(c, hb, hf, 0, o);
}

// One mixed step with the learned field step attached (direct launches):
// capture mode appends the ADI's (E, S, E_next) pairs of every fno_every-th
// slice to the network's training set; replace mode advances the field with
// the network and runs the ADI instead on every fno_every-th step and when
// the energy-balance residual of the network's step exceeds fno_thr.
static int mix_step_fno(Ctx *c, const MixKey &k) {
    const MixArrays a = mix_arrays(c);
    const int tm = std::get<0>(k), maxzero = std::get<4>(k);
    const bool refine = std::get<1>(k) != 0, sc = std::get<2>(k) != 0, moments = std::get<3>(k) != 0;
    const int adi_mode = std::get<5>(k);
    const int ns = c->ns;
    const long long P = (long long)c->ng * c->ng;
    int ntiles = 0;
    MixState hs{};
    if (c->fno_mode == 1) {
        launch_mix_beam(c->stream, a, sc, moments, nullptr, false, std::get<6>(k), std::get<7>(k),
                        std::get<8>(k) != 0);
        if (launch_mix_adi(c->stream, a, tm, refine, adi_mode, &ntiles)) {
            set_error("mixed form: no ADI instantiation for this band width");
            return G4MI_E_ARG;
        }
        std::vector<int4> box(ns);
        G4_HIP(hipMemcpyAsync(&hs, c->m_st, sizeof(MixState), hipMemcpyDeviceToHost, c->stream));
        G4_HIP(hipMemcpyAsync(box.data(), c->d_box, ns * sizeof(int4), hipMemcpyDeviceToHost, c->stream));
        G4_HIP(hipStreamSynchronize(c->stream));
        for (int ii = 0; ii < ns; ii += c->fno_every) {
            const long long fs = (ii + hs.first) % ns;
            if (fno_capture(c->fno, c->stream, c->m_field[hs.cur] + fs * P, c->m_src + (long long)ii * P,
                            c->m_field[hs.cur ^ 1] + fs * P, box[ii]) != G4MI_OK)
                break;  // training set full
        }
        launch_mix_tail(c->stream, a, maxzero, ntiles, nullptr);
        G4_HIP(hipGetLastError());
        return G4MI_OK;
    }
    // replace mode; after kFnoRejectRun residual rejections in a row the
    // next kFnoSkip steps run the ADI alone (no inference to pay for)
    constexpr int kFnoRejectRun = 4, kFnoSkip = 8;
    {
        const long long stepno0 = c->fno_steps + c->fno_adi_steps;
        const bool judge0 = c->fno_every > 0 && stepno0 % c->fno_every == 0;
        if (c->fno_skip > 0 && !judge0) {
            c->fno_skip--;
            c->fno_skipped++;
            c->fno_adi_steps++;
            launch_mix_beam(c->stream, a, sc, moments, nullptr, false, std::get<6>(k), std::get<7>(k),
                        std::get<8>(k) != 0);
            if (launch_mix_adi(c->stream, a, tm, refine, adi_mode, &ntiles)) {
                set_error("mixed form: no ADI instantiation for this band width");
                return G4MI_E_ARG;
            }
            launch_mix_tail(c->stream, a, maxzero, ntiles, nullptr);
            G4_HIP(hipGetLastError());
            return G4MI_OK;
        }
    }
    const double ks = 4. * asin(1) / c->xlambda;
    const double scl = c->dgrid * kEev / ks, pw = scl * scl / kVacimp;
    std::vector<double> hf0(8 * (size_t)ns), hf1(8 * (size_t)ns), eg(ns), ss(ns);
    launch_mix_field_sums(c->stream, a, 0, c->d_diag_f);
    G4_HIP(hipMemcpyAsync(hf0.data(), c->d_diag_f, hf0.size() * 8, hipMemcpyDeviceToHost, c->stream));
    launch_mix_beam(c->stream, a, sc, moments, nullptr, true, std::get<6>(k), std::get<7>(k),
                        std::get<8>(k) != 0);
    G4_HIP(hipMemcpyAsync(&hs, c->m_st, sizeof(MixState), hipMemcpyDeviceToHost, c->stream));
    G4_HIP(hipStreamSynchronize(c->stream));
    G4_TRY(fno_infer_slices(c->fno, c->stream, c->m_field[hs.cur], c->m_field[hs.cur ^ 1], c->m_src, c->d_box, ns,
                            hs.first));
    launch_mix_field_sums(c->stream, a, 1, c->d_diag_f);
    G4_HIP(hipMemcpyAsync(hf1.data(), c->d_diag_f, hf1.size() * 8, hipMemcpyDeviceToHost, c->stream));
    G4_HIP(hipMemcpyAsync(eg.data(), c->m_egsum, eg.size() * 8, hipMemcpyDeviceToHost, c->stream));
    launch_mix_src_sums(c->stream, a, c->d_diag_b);
    G4_HIP(hipMemcpyAsync(ss.data(), c->d_diag_b, ss.size() * 8, hipMemcpyDeviceToHost, c->stream));
    G4_HIP(hipStreamSynchronize(c->stream));
    // particle-to-field energy balance of the step.  A step's field change is
    // dP = 2 Re<U E, V S> + |V S|^2 (U unitary, V ~ 2 to first order in the
    // step): the beam's loss sum_s I_s eev d<gamma>_s (from the push's own
    // increments, free of the storage rounding) balances the cross term, and
    // the source's self-energy |2 S|^2 is subtracted.
    double p0 = 0, p1 = 0, dpb = 0, pself = 0;
    for (int i = 0; i < ns; i++) {
        p0 += hf0[8 * (size_t)i] * pw;
        p1 += hf1[8 * (size_t)i] * pw;
        pself += 4.0 * ss[i] * pw;
        const long long n = c->off[i + 1] - c->off[i];
        if (n > 0) dpb += c->h_cur[i] * kEev * eg[i] / (double)n;
    }
    const double res = std::fabs(p1 - p0 + dpb - pself) / std::max(p1, 1e-300);
    c->fno_res_all_sum += res;
    c->fno_nres_all++;
    const long long stepno = c->fno_steps + c->fno_adi_steps;
    const bool judge = c->fno_every > 0 && stepno % c->fno_every == 0;
    const bool use_adi = judge || !(res <= c->fno_thr);
    if (!judge) {
        if (use_adi) {
            if (++c->fno_reject_run >= kFnoRejectRun) {
                c->fno_skip = kFnoSkip;
                c->fno_reject_run = 0;
            }
        } else {
            c->fno_reject_run = 0;
        }
    }
    if (use_adi) {
        if (judge) {
            if (!c->m_judge) {
                G4_TRY(dalloc(c, &c->m_judge, (size_t)ns * P));
                G4_TRY(dalloc(c, &c->d_judge, (size_t)ns * 3));
            }
            G4_HIP(hipMemcpyAsync(c->m_judge, c->m_field[hs.cur ^ 1], (size_t)ns * P * sizeof(float2),
                                  hipMemcpyDeviceToDevice, c->stream));
        }
        if (launch_mix_adi(c->stream, a, tm, refine, adi_mode, &ntiles)) {
            set_error("mixed form: no ADI instantiation for this band width");
            return G4MI_E_ARG;
        }
        if (judge) {
            // the network's step against the ADI's, over every slice
            fno_judge_sums(c->stream, c->m_judge, c->m_field[hs.cur ^ 1], c->m_field[hs.cur], ns, (int)P,
                           c->d_judge);
            std::vector<double> js(3 * (size_t)ns);
            G4_HIP(hipMemcpyAsync(js.data(), c->d_judge, js.size() * 8, hipMemcpyDeviceToHost, c->stream));
            G4_HIP(hipStreamSynchronize(c->stream));
            double e = 0, u = 0, t = 0;
            for (int j = 0; j < ns; j++) {
                e += js[3 * (size_t)j];
                u += js[3 * (size_t)j + 1];
                t += js[3 * (size_t)j + 2];
            }
            if (t > 0) c->fno_judge_field += std::sqrt(e / t);
            if (u > 0) c->fno_judge_update += std::sqrt(e / u);
            c->fno_njudge++;
        }
        // the ADI's own residual, for reference
        std::vector<double> hfa(8 * (size_t)ns);
        launch_mix_field_sums(c->stream, a, 1, c->d_diag_f);
        G4_HIP(hipMemcpyAsync(hfa.data(), c->d_diag_f, hfa.size() * 8, hipMemcpyDeviceToHost, c->stream));
        G4_HIP(hipStreamSynchronize(c->stream));
        double pa = 0;
        for (int i = 0; i < ns; i++) pa += hfa[8 * (size_t)i] * pw;
        c->fno_res_adi_sum += std::fabs(pa - p0 + dpb - pself) / std::max(pa, 1e-300);
        c->fno_nres_adi++;
        c->fno_adi_steps++;
    } else {
        launch_mix_field_sums(c->stream, a, 1, c->m_tilediag);
        ntiles = 1;
        c->fno_steps++;
        c->fno_res_sum += res;
    }
    launch_mix_tail(c->stream, a, maxzero, ntiles, nullptr);
    G4_HIP(hipGetLastError());
    return G4MI_OK;
}

// One step of the mixed-form schedule tuner (see mix_run): tuning step k
// runs main candidate k % 8 with ADI candidate (k / 8 + k) % 3 (k < 16: the
// whole-width streaming kernel, the strips, the tiles), the graph of the
// chosen kernels (k = 16) or the persistent launch (k = 17).  The main
// candidates: 6, 4 or 8 wavefronts per slice with 68 or 44 KiB of LDS, and
// 4 or 6 wavefronts with 32 KiB (more workgroups per CU, a smaller
// deposit tile).
constexpr int kMixTuneSteps = 18;
constexpr int kMainCands = 8;
static const int kMainCand[kMainCands][2] = {{384, 0}, {256, 0}, {512, 0}, {384, 1},
                                              {256, 1}, {512, 1}, {256, 2}, {384, 2}};

static int mix_tune_step(Ctx *c, const MixKey &key) {
    hipEvent_t *tev = c->m_tune_ev;
    const int k = c->m_tune_i;
    if (k == 0)
        for (int e = 0; e < 9; e++) G4_HIP(hipEventCreate(&tev[e]));
    MixKey kk = key;
    float ms = 0;
    if (k < 2 * kMainCands) {
        // every main candidate twice, each ADI kernel four to six times
        const int mc = k % kMainCands, ac = (k / kMainCands + k) % 3;
        const int tm = std::get<0>(kk);
        const bool tile_ok = !c->m_refine || tm <= 9;
        const int mode = ac == 0 ? kAdiWide : ac == 1 ? kAdiStrips : (tile_ok ? kAdiTile : kAdiWide);
        std::get<1>(kk) = c->m_refine ? 1 : 0;
        std::get<5>(kk) = mode;
        std::get<6>(kk) = kMainCand[mc][0];
        std::get<7>(kk) = kMainCand[mc][1];
        G4_TRY(mix_launch(c, kk, tev));
        G4_HIP(hipEventSynchronize(tev[8]));
        float tm_main = 0, tm_adi = 0;
        G4_HIP(hipEventElapsedTime(&tm_main, tev[2], tev[3]));
        G4_HIP(hipEventElapsedTime(&tm_adi, tev[6], tev[7]));
        c->m_tune_main[mc] += tm_main;
        c->m_tune_nmain[mc]++;
        c->m_tune_adi[mode] += tm_adi;
        c->m_tune_nadi[mode]++;
        if (k == 2 * kMainCands - 1) {
            int best = 0;
            for (int q = 1; q < kMainCands; q++)
                if (c->m_tune_nmain[q] &&
                    c->m_tune_main[q] / c->m_tune_nmain[q] < c->m_tune_main[best] / c->m_tune_nmain[best])
                    best = q;
            c->m_main_nt = kMainCand[best][0];
            c->m_main_lds = kMainCand[best][1];
            int ba = -1;
            for (int q = 0; q < 3; q++)
                if (c->m_tune_nadi[q] &&
                    (ba < 0 || c->m_tune_adi[q] / c->m_tune_nadi[q] < c->m_tune_adi[ba] / c->m_tune_nadi[ba]))
                    ba = q;
            c->m_adi_mode = ba < 0 ? kAdiWide : ba;
        }
    } else {
        // the chosen kernels for this key
        const int tm = std::get<0>(kk);
        std::get<1>(kk) = (c->m_refine && (c->m_adi_mode != kAdiTile || tm <= 9)) ? 1 : 0;
        std::get<5>(kk) = c->m_adi_mode;
        std::get<6>(kk) = c->m_main_nt;
        std::get<7>(kk) = c->m_main_lds;
        if (k == 2 * kMainCands) {
            hipGraphExec_t ex = nullptr;
            G4_TRY(mix_graph(c, kk, &ex));
            G4_HIP(hipEventRecord(tev[0], c->stream));
            G4_HIP(hipGraphLaunch(ex, c->stream));
            G4_HIP(hipEventRecord(tev[8], c->stream));
            G4_HIP(hipEventSynchronize(tev[8]));
            G4_HIP(hipEventElapsedTime(&ms, tev[0], tev[8]));
            c->m_tune_graph = ms;
        } else {
            double pers = -1;
            if (c->m_adi_mode != kAdiTile && launch_mix_step_persistent(c->stream, mix_arrays(c), tm, std::get<1>(kk) != 0,
                                                          std::get<2>(kk) != 0, std::get<3>(kk) != 0,
                                                          std::get<4>(kk), tev) == 0) {
                G4_HIP(hipEventSynchronize(tev[8]));
                G4_HIP(hipEventElapsedTime(&ms, tev[0], tev[8]));
                pers = ms;
            } else {
                (void)hipGetLastError();  // refused launch: run the step as kernels
                G4_TRY(mix_launch(c, kk, nullptr));
            }
            c->m_persistent = pers > 0 && pers < c->m_tune_graph;
            char buf[768];
            const double *mm = c->m_tune_main;
            const int *nm = c->m_tune_nmain;
            auto avg = [](double s, int n) { return n ? s / n : -1.0; };
            snprintf(buf, sizeof(buf),
                     "{\"main_ms\": {\"384/68K\": %.4f, \"256/68K\": %.4f, \"512/68K\": %.4f, "
                     "\"384/44K\": %.4f, \"256/44K\": %.4f, \"512/44K\": %.4f, \"256/32K\": %.4f, "
                     "\"384/32K\": %.4f}, "
                     "\"adi_ms\": {\"wide\": %.4f, \"strips\": %.4f, \"tile\": %.4f}, "
                     "\"step_ms\": {\"graph\": %.4f, \"persistent\": %.4f}}",
                     avg(mm[0], nm[0]), avg(mm[1], nm[1]), avg(mm[2], nm[2]), avg(mm[3], nm[3]),
                     avg(mm[4], nm[4]), avg(mm[5], nm[5]), avg(mm[6], nm[6]), avg(mm[7], nm[7]),
                     avg(c->m_tune_adi[kAdiWide], c->m_tune_nadi[kAdiWide]),
                     avg(c->m_tune_adi[kAdiStrips], c->m_tune_nadi[kAdiStrips]),
                     avg(c->m_tune_adi[kAdiTile], c->m_tune_nadi[kAdiTile]), c->m_tune_graph, pers);
            c->m_tune_log = buf;
            c->m_tuned = true;
            for (int e = 0; e < 9; e++) (void)hipEventDestroy(tev[e]);
        }
    }
    c->m_tune_i = k + 1;
    return G4MI_OK;
}

static int mix_run(Ctx *c, const g4mi_step_params *steps, int nsteps, double gamma0,
                   g4mi_diag_out *records, int *done) {
    *done = 0;
    if (!c->has_beam || !c->has_field) {
        set_error("beam and field must be uploaded before stepping");
        return G4MI_E_STATE;
    }
    int n = 0;
    while (n < nsteps && !(steps[n].marker & 8)) n++;
    if (n == 0) return G4MI_OK;
    std::vector<MixStep> tab(n);
    std::vector<MixKey> keys(n);
    const bool sc = (c->sc_nz > 0 && c->sc_ng > 2);
    const bool moments = sc || c->longrange;
    int maxzero = 0;
    if (c->timerun && !c->periodic)
        for (int i = 0; i < n; i++)
            maxzero = std::max(maxzero, (int)std::ceil(std::fabs(steps[i].slip) / (0.8 * c->sample)) + 1);
    if (sc && (2 * c->sc_nphi + 1) * c->sc_nz * c->sc_ng > mix_sc_mode_cells()) {
        set_error("mixed form: (2 nphi + 1) nz x radial cells of the space-charge solver exceed " +
                  std::to_string(mix_sc_mode_cells()));
        return G4MI_E_ARG;
    }
    if (maxzero > kMixMaxZero) {
        set_error("mixed form: a step slips more slices than the device zero list holds");
        return G4MI_E_ARG;
    }
    for (int i = 0; i < n; i++) {
        G4_TRY(mix_fill_step(c, steps[i], gamma0, tab[i]));
        const int tm = c->band_tmpl[tab[i].band];
        // the tile kernel's correction needs more LDS than a CU has beyond a half width of 9
        const int refine = (c->m_refine && (c->m_adi_mode != kAdiTile || tm <= 9)) ? 1 : 0;
        keys[i] = MixKey(tm, refine, sc ? 1 : 0, moments ? 1 : 0, maxzero, c->m_adi_mode, c->m_main_nt,
                         c->m_main_lds, 1);
    }
    // the main kernel of every step but the run's last also runs the next
    // step's front, while its slice's particles are fresh in the caches; the
    // particles are back in the reference's order between runs
    if (c->m_fuse_front)
        for (int i = 0; i + 1 < n; i++) {
            tab[i].fnext = tab[i + 1].front_done = 1;
            tab[i + 1].fbuf = tab[i].fbuf ^ 1;
        }
    for (int i = 0; i < n; i++) std::get<8>(keys[i]) = tab[i].front_done ? 0 : 1;
    if (n > c->m_steps_cap) {
        if (c->m_steps) (void)hipFree(c->m_steps);
        c->m_steps = nullptr;
        c->m_steps_cap = 0;
        G4_HIP(hipMalloc(&c->m_steps, (size_t)std::max(n, 1024) * sizeof(MixStep)));
        c->m_steps_cap = std::max(n, 1024);
        mix_clear_graphs(c);  // the table pointer is part of every captured launch
    }
    G4_HIP(hipMemcpyAsync(c->m_steps, tab.data(), (size_t)n * sizeof(MixStep), hipMemcpyHostToDevice,
                          c->stream));
    const int zero = 0;
    G4_HIP(hipMemcpyAsync(&c->m_st->istep, &zero, sizeof(int), hipMemcpyHostToDevice, c->stream));
    G4_HIP(hipStreamSynchronize(c->stream));
    const size_t rowd = (size_t)c->ns * kMixRecord;
    // Records: the device ring of m_nrec rows (row = step mod m_nrec) goes to
    // its pinned host mirror half by half, each half by an async copy queued
    // behind its steps; the host finishes a half's rows just before the next
    // copy into that half (its own copy ended a half's worth of steps
    // earlier) or at the end of the run, so the stream never waits on the host.
    const int half = c->m_nrec / 2;
    int pending = 0;                       // first step whose record copy is not queued
    int hfirst[2] = {0, 0}, hcnt[2] = {0, 0};
    auto finish_half = [&](int h) -> int {
        if (hcnt[h] == 0) return G4MI_OK;
        G4_HIP(hipEventSynchronize(c->m_rec_ev[h]));
        for (int k = 0; k < hcnt[h]; k++) {
            const int stp = hfirst[h] + k;
            mix_finish_row(c, c->m_rec_pin + (size_t)(stp % c->m_nrec) * rowd, &records[stp]);
        }
        hcnt[h] = 0;
        return G4MI_OK;
    };
    // The schedule autotuner: the first kMixTuneSteps steps of a context
    // (over as many g4mi_run calls as they take) each run one candidate of the
    // main kernel's shape (threads per slice x LDS region) and one of the ADI
    // kernels (streaming strips or 32 x 32 tiles), timed by HIP events; the
    // fastest of each is kept; two more steps then time the chosen kernels as
    // one replayed graph and the whole step as one persistent cooperative
    // launch.  Every candidate computes the same step, so the tuning steps
    // are steps of the run, not repeats.
    for (int i = 0; i < n; i++) {
        const bool tune = c->m_autotune && !c->m_tuned && !(c->fno && c->fno_mode) && !c->timing;
        if (tune) {
            G4_TRY(mix_tune_step(c, keys[i]));
            if (c->m_tuned)
                for (int j = i + 1; j < n; j++) {
                    const int tmj = std::get<0>(keys[j]);
                    std::get<1>(keys[j]) = (c->m_refine && (c->m_adi_mode != kAdiTile || tmj <= 9)) ? 1 : 0;
                    std::get<5>(keys[j]) = c->m_adi_mode;
                    std::get<6>(keys[j]) = c->m_main_nt;
                    std::get<7>(keys[j]) = c->m_main_lds;
                }
        } else if (c->fno && c->fno_mode) {
            G4_TRY(mix_step_fno(c, keys[i]));
        } else if (c->m_persistent && !c->timing) {
            const MixKey &k = keys[i];
            if (launch_mix_step_persistent(c->stream, mix_arrays(c), std::get<0>(k), std::get<1>(k) != 0,
                                           std::get<2>(k) != 0, std::get<3>(k) != 0, std::get<4>(k), nullptr)) {
                set_error("mixed form: the persistent step launch was refused");
                return G4MI_E_HIP;
            }
        } else if (c->timing) {
            // timed step: direct launches with the kernel groups bracketed by events
            G4_TRY(step_begin_timing(c));
            G4_TRY(mix_launch(c, keys[i], c->ev));
            c->ev = nullptr;
            c->pool_used++;
        } else if (c->m_graphs) {
            hipGraphExec_t ex = nullptr;
            G4_TRY(mix_graph(c, keys[i], &ex));
            G4_HIP(hipGraphLaunch(ex, c->stream));
        } else {
            G4_TRY(mix_launch(c, keys[i], nullptr));
        }
        if (records && ((i + 1 - pending) == half || i == n - 1)) {
            // steps pending..i: whole halves but the last, so their rows are
            // contiguous within one half of the ring
            const int cnt = i + 1 - pending, h = (pending / half) & 1;
            G4_TRY(finish_half(h));
            const size_t r0 = (size_t)(pending % c->m_nrec) * rowd;
            G4_HIP(hipMemcpyAsync(c->m_rec_pin + r0, c->m_rec + r0, (size_t)cnt * rowd * sizeof(double),
                                  hipMemcpyDeviceToHost, c->stream));
            G4_HIP(hipEventRecord(c->m_rec_ev[h], c->stream));
            hfirst[h] = pending;
            hcnt[h] = cnt;
            pending = i + 1;
        }
    }
    G4_TRY(finish_half(0));
    G4_TRY(finish_half(1));
    MixState hs{};
    G4_HIP(hipMemcpyAsync(&hs, c->m_st, sizeof(MixState), hipMemcpyDeviceToHost, c->stream));
    G4_HIP(hipStreamSynchronize(c->stream));
    c->first = hs.first;
    c->accuslip = hs.accuslip;
    *done = n;
    return G4MI_OK;
}

static int mix_diag(Ctx *c, g4mi_diag_out *o) {
    const MixArrays a = mix_arrays(c);
    launch_mix_diag(c->stream, a, c->d_diag_b, c->d_diag_f);
    const int ns = c->ns;
    std::vector<double> hb(8 * (size_t)ns), hf(8 * (size_t)ns);
    MixState hs{};
    G4_HIP(hipMemcpyAsync(hb.data(), c->d_diag_b, hb.size() * 8, hipMemcpyDeviceToHost, c->stream));
    G4_HIP(hipMemcpyAsync(hf.data(), c->d_diag_f, hf.size() * 8, hipMemcpyDeviceToHost, c->stream));
    G4_HIP(hipMemcpyAsync(&hs, c->m_st, sizeof(MixState), hipMemcpyDeviceToHost, c->stream));
    G4_HIP(hipStreamSynchronize(c->stream));
    finish_diag(c, hb, hf, hs.first, o);
    if (wants_aux(o)) {
        std::vector<double> aux, esc, ssc;
        G4_TRY(fetch_aux(c, aux, esc, ssc));
        finish_aux(c, aux.data(), esc.data(), ssc.data(), c->wake_applied, o);
    }
    if (wants_spectral(o)) {  // on an FP64 copy of the FP32 field
        G4_TRY(ensure_dft(c));
        launch_widen_field(c->stream, c->m_field[hs.cur], c->d_fft_u, (long long)ns * c->ng * c->ng);
        G4_TRY(spectral_diag(c, c->d_fft_u, hs.first, o));
    }
    return G4MI_OK;
}

static int mix_split_call(const char *name) {
    set_error(std::string(name) + ": the mixed form advances whole steps (g4mi_step, g4mi_run)");
    return G4MI_E_STATE;
}

// Upload in offset form: gref = the beam's mean energy rounded to an FP32
// value, base = 2 pi floor(<theta>_slice / 2 pi) per slice.
static int mix_beam_upload(Ctx *c, const double *g, const double *th, const double *x,
                           const double *y, const double *px, const double *py, const double *cur) {
    const long long n = c->ntot;
    const int ns = c->ns;
    double gsum = 0;
    for (long long i = 0; i < n; i++) gsum += g[i];
    c->gref = n > 0 ? (double)(float)(gsum / (double)n) : 1.0;
    const double twopi = 4. * asin(1.0);
    std::vector<double> base(ns, 0.0);
    std::vector<float> hv(6 * (size_t)n);
    float *fg = hv.data(), *ft = fg + n, *fx = ft + n, *fy = fx + n, *fpx = fy + n, *fpy = fpx + n;
    for (int is = 0; is < ns; is++) {
        const long long a = c->off[is], b = c->off[is + 1];
        double ts = 0;
        for (long long i = a; i < b; i++) ts += th[i];
        if (b > a) base[is] = twopi * floor(ts / (double)(b - a) / twopi);
        for (long long i = a; i < b; i++) {
            fg[i] = (float)(g[i] - c->gref);
            ft[i] = (float)(th[i] - base[is]);
            fx[i] = (float)x[i];
            fy[i] = (float)y[i];
            fpx[i] = (float)px[i];
            fpy[i] = (float)py[i];
        }
    }
    float *dst[6] = {c->m_dg, c->m_dth, c->m_x, c->m_y, c->m_px, c->m_py};
    for (int k = 0; k < 6; k++)
        if (n) G4_HIP(hipMemcpyAsync(dst[k], hv.data() + (size_t)k * n, (size_t)n * sizeof(float),
                                     hipMemcpyHostToDevice, c->stream));
    G4_HIP(hipMemcpyAsync(c->m_base, base.data(), ns * sizeof(double), hipMemcpyHostToDevice, c->stream));
    G4_HIP(hipMemcpyAsync(c->d_cur, cur, ns * sizeof(double), hipMemcpyHostToDevice, c->stream));
    G4_HIP(hipStreamSynchronize(c->stream));
    c->h_cur.assign(cur, cur + ns);
    mix_clear_graphs(c);  // gref is part of every captured launch
    c->has_beam = true;
    return G4MI_OK;
}

static int mix_beam_download(Ctx *c, double *g, double *th, double *x, double *y, double *px,
                             double *py) {
    const long long n = c->ntot;
    const int ns = c->ns;
    std::vector<float> hv(6 * (size_t)n);
    std::vector<double> base(ns);
    const float *src[6] = {c->m_dg, c->m_dth, c->m_x, c->m_y, c->m_px, c->m_py};
    for (int k = 0; k < 6; k++)
        if (n) G4_HIP(hipMemcpyAsync(hv.data() + (size_t)k * n, src[k], (size_t)n * sizeof(float),
                                     hipMemcpyDeviceToHost, c->stream));
    G4_HIP(hipMemcpyAsync(base.data(), c->m_base, ns * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    G4_HIP(hipStreamSynchronize(c->stream));
    const float *fg = hv.data(), *ft = fg + n;
    for (int is = 0; is < ns; is++)
        for (long long i = c->off[is]; i < c->off[is + 1]; i++) {
            if (g) g[i] = c->gref + (double)fg[i];
            if (th) th[i] = base[is] + (double)ft[i];
        }
    double *dst[4] = {x, y, px, py};
    for (int k = 0; k < 4; k++)
        if (dst[k])
            for (long long i = 0; i < n; i++) dst[k][i] = (double)hv[(size_t)(k + 2) * n + i];
    return G4MI_OK;
}

static int mix_field_upload(Ctx *c, const double *field, int first) {
    const size_t cells = (size_t)c->ns * c->ng * c->ng;
    std::vector<float2> hf(cells);
    for (size_t k = 0; k < cells; k++) hf[k] = make_float2((float)field[2 * k], (float)field[2 * k + 1]);
    MixState hs{};
    G4_HIP(hipMemcpyAsync(&hs, c->m_st, sizeof(MixState), hipMemcpyDeviceToHost, c->stream));
    G4_HIP(hipStreamSynchronize(c->stream));
    hs.first = first;
    hs.cur = 0;
    hs.nzero = 0;
    G4_HIP(hipMemcpyAsync(c->m_field[0], hf.data(), cells * sizeof(float2), hipMemcpyHostToDevice, c->stream));
    G4_HIP(hipMemcpyAsync(c->m_st, &hs, sizeof(MixState), hipMemcpyHostToDevice, c->stream));
    G4_HIP(hipStreamSynchronize(c->stream));
    c->first = first;
    c->has_field = true;
    return G4MI_OK;
}

static int mix_field_download(Ctx *c, double *field, int *first) {
    MixState hs{};
    G4_HIP(hipMemcpyAsync(&hs, c->m_st, sizeof(MixState), hipMemcpyDeviceToHost, c->stream));
    G4_HIP(hipStreamSynchronize(c->stream));
    if (field) {
        const size_t cells = (size_t)c->ns * c->ng * c->ng;
        std::vector<float2> hf(cells);
        G4_HIP(hipMemcpyAsync(hf.data(), c->m_field[hs.cur], cells * sizeof(float2), hipMemcpyDeviceToHost,
                              c->stream));
        G4_HIP(hipStreamSynchronize(c->stream));
        for (size_t k = 0; k < cells; k++) {
            field[2 * k] = hf[k].x;
            field[2 * k + 1] = hf[k].y;
        }
    }
    if (first) *first = hs.first;
    return G4MI_OK;
}

static int mix_alloc(Ctx *c) {
    const size_t n = (size_t)c->ntot, ns = (size_t)c->ns, cells = (size_t)c->ng * c->ng;
    int rc;
    if ((rc = dalloc(c, &c->m_dg, n)) || (rc = dalloc(c, &c->m_dth, n)) || (rc = dalloc(c, &c->m_x, n)) ||
        (rc = dalloc(c, &c->m_y, n)) || (rc = dalloc(c, &c->m_px, n)) || (rc = dalloc(c, &c->m_py, n)) ||
        (rc = dalloc(c, &c->m_xs, n)) || (rc = dalloc(c, &c->m_ys, n)) ||
        (rc = dalloc(c, &c->m_base, ns)) || (rc = dalloc(c, &c->m_pbox, ns)) ||
        (rc = dalloc(c, &c->m_field[0], ns * cells)) ||
        (rc = dalloc(c, &c->m_field[1], ns * cells)) || (rc = dalloc(c, &c->m_src, ns * cells)) ||
        (rc = dalloc(c, &c->m_beamsum, 8 * ns)) || (rc = dalloc(c, &c->m_egsum, ns)) ||
        (rc = dalloc(c, &c->m_mom_next, ns)) ||
        (rc = dalloc(c, &c->m_tilediag, ns * mix_band_tiles(c->ng) * 8)) ||
        (rc = dalloc(c, &c->m_rec, (size_t)kMixRecRows * ns * kMixRecord)) ||
        (rc = dalloc(c, &c->m_st, 1)) ||
        (rc = dalloc(c, &c->m_bands, (size_t)kMixMaxBands * c->ng * kMixBandRowMax)) ||
        (rc = dalloc(c, &c->m_binfo, (size_t)kMixMaxBands)))
        return rc;
    c->m_nrec = kMixRecRows;
    G4_HIP(hipHostMalloc((void **)&c->m_rec_pin, (size_t)kMixRecRows * ns * kMixRecord * sizeof(double)));
    for (auto &e : c->m_rec_ev) G4_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    if (mix_prepare() != 0) {
        set_error("mixed form: setting the kernels' LDS limits failed");
        return G4MI_E_HIP;
    }
    MixState hs{};
    G4_HIP(hipMemcpy(c->m_st, &hs, sizeof(MixState), hipMemcpyHostToDevice));
    return G4MI_OK;
}

}  // namespace g4mi

namespace g4mi {
// ---------------------------------------------------------------------------
// Scratch of the allocation-free entry point (g4mi_sort_index_noalloc): a
// file-scope device array, one copy per GPU at its own address, held until
// the process ends.  Layout: the moved-particle counter (8 bytes) at 0, the
// target slice of every particle (int) from kScratchAtar on.  Host code
// never passes `scratch` itself to a kernel (on the host it names a
// host-side stand-in): it uses the per-device address that
// hipGetSymbolAddress returns, taken once per device at context creation.
// ---------------------------------------------------------------------------
constexpr long long kNoAllocMaxParticles = 1LL << 22;   // 4194304 particles: 16 MiB of targets
constexpr size_t kScratchAtar = 256;
constexpr size_t kScratchBytes = kScratchAtar + (size_t)kNoAllocMaxParticles * sizeof(int);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];          // device address of `scratch`, per device
static hipError_t g_scratch_err[kMaxDevices];     // hipGetSymbolAddress status, per device

// Once per device, with that device current.
static void scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_err[device] = e;
    if (e == hipSuccess) g_scratch_ptr[device] = p;
}

__global__ void k_clear_u64(unsigned long long *p) {
    if (threadIdx.x == 0) *p = 0ULL;
}
}  // namespace g4mi

using namespace g4mi;

struct g4mi_ctx {
    Ctx c;
};

extern "C" {

const char *g4mi_last_error(void) { return g_last_error.c_str(); }

int g4mi_create(g4mi_ctx **out, int device, int nslice, const long long *off, int ngrid,
                double gridmax, double lambda0, double sample, int timerun, int periodic,
                int precision) {
    if (!out || nslice < 1 || !off || ngrid < 3 || gridmax <= 0 || lambda0 <= 0 || sample <= 0) {
        set_error("g4mi_create: invalid argument");
        return G4MI_E_ARG;
    }
    if (ngrid > 4096) {
        set_error("g4mi_create: ngrid above 4096 is not supported");
        return G4MI_E_ARG;
    }
    if (precision != G4MI_FP64 && precision != G4MI_FP32) {
        set_error("g4mi_create: precision must be G4MI_FP64 or G4MI_FP32");
        return G4MI_E_ARG;
    }
    for (int i = 0; i < nslice; i++)
        if (off[i + 1] < off[i] || off[0] != 0) {
            set_error("g4mi_create: slice offsets must start at 0 and be non-decreasing");
            return G4MI_E_ARG;
        }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1 || device < 0 || device >= ndev) {
        set_error("g4mi_create: no usable HIP device");
        return G4MI_E_NODEV;
    }
    G4_HIP(hipSetDevice(device));
    scratch_setup(device);  // the allocation-free entry point's scratch address on this device
    g4mi_ctx *h = new g4mi_ctx;
    Ctx *c = &h->c;
    *out = nullptr;
    c->device = device;
    c->precision = precision;
    c->ns = nslice;
    c->ng = ngrid;
    c->gridmax = gridmax;
    c->dgrid = 2 * gridmax / (double)(ngrid - 1);  // Field::init :45-46
    c->lambda0 = lambda0;
    c->xlambda = lambda0 / 1.0;
    c->xks = 4. * asin(1) / c->xlambda;
    c->sample = sample;
    c->slicelength = sample * lambda0;
    c->timerun = timerun;
    c->periodic = periodic;
    c->off.assign(off, off + nslice + 1);
    c->ntot = off[nslice];
    for (int i = 0; i < nslice; i++) c->maxn = std::max(c->maxn, off[i + 1] - off[i]);
    int rc = G4MI_OK;
    auto fail = [&](int r) {
        g4mi_destroy(h);
        return r;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return fail(G4MI_E_HIP);
    const bool fp64 = precision == G4MI_FP64;
    // the FP64 particle and field arrays exist only in the plain form
    const size_t n = fp64 ? (size_t)c->ntot : 0, ns = (size_t)nslice,
                 cells = fp64 ? (size_t)ngrid * ngrid : 0;
    if ((rc = dalloc(c, &c->d_off, ns + 1)) || (rc = dalloc(c, &c->d_g, n)) ||
        (rc = dalloc(c, &c->d_th, n)) || (rc = dalloc(c, &c->d_x, n)) ||
        (rc = dalloc(c, &c->d_y, n)) || (rc = dalloc(c, &c->d_px, n)) ||
        (rc = dalloc(c, &c->d_py, n)) || (rc = dalloc(c, &c->d_ez, n)) ||
        (rc = dalloc(c, &c->d_cur, ns)) || (rc = dalloc(c, &c->d_esc, ns)) ||
        (rc = dalloc(c, &c->d_mom, 4 * ns)) || (rc = dalloc(c, &c->d_fsize, ns)) ||
        (rc = dalloc(c, &c->d_cen, 2 * ns)) || (rc = dalloc(c, &c->d_rbound, ns)) ||
        (rc = dalloc(c, &c->d_dr, ns)) || (rc = dalloc(c, &c->d_rmax, 1)) ||
        (rc = dalloc(c, &c->d_efield, ns)) || (rc = dalloc(c, &c->d_field, ns * cells)) ||
        (rc = dalloc(c, &c->d_work, ns * cells)) || (rc = dalloc(c, &c->d_src, ns * cells)) ||
        (rc = dalloc(c, &c->d_box, ns)) || (rc = dalloc(c, &c->d_fac, 3 * (size_t)ngrid)) ||
        (rc = dalloc(c, &c->d_diag_b, 8 * ns)) || (rc = dalloc(c, &c->d_diag_f, 8 * ns)) ||
        (rc = dalloc(c, &c->d_count, 1)) || (rc = dalloc(c, &c->d_beamsum, 8 * ns)) ||
        (rc = dalloc(c, &c->d_pbox, ns)) || (rc = dalloc(c, &c->d_pmaxf, ns)))
        return fail(rc);
    if (!fp64 && (rc = mix_alloc(c))) return fail(rc);
    if (fp64 && (rc = alloc_records(c))) return fail(rc);
    std::vector<int> bs;
    std::vector<long long> bb;
    build_block_table(c->off, bs, bb);
    c->nblk = (int)bs.size();
    // capacity for any redistribution of the particles over the slices
    c->blk_cap = (int)(ns + (size_t)c->ntot / 256 + 1);
    if ((rc = dalloc(c, &c->d_blk_slice, (size_t)c->blk_cap)) || (rc = dalloc(c, &c->d_blk_begin, (size_t)c->blk_cap)))
        return fail(rc);
    if (hipMemcpy(c->d_off, off, (ns + 1) * sizeof(long long), hipMemcpyHostToDevice) != hipSuccess ||
        (c->nblk && hipMemcpy(c->d_blk_slice, bs.data(), bs.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) ||
        (c->nblk && hipMemcpy(c->d_blk_begin, bb.data(), bb.size() * sizeof(long long), hipMemcpyHostToDevice) != hipSuccess) ||
        hipMemset(c->d_esc, 0, ns * sizeof(double)) != hipSuccess ||
        hipMemset(c->d_ez, 0, (n ? n : 1) * sizeof(double)) != hipSuccess ||
        hipMemset(c->d_efield, 0, ns * sizeof(double)) != hipSuccess ||
        hipMemset(c->d_rmax, 0, sizeof(double)) != hipSuccess) {
        set_error("g4mi_create: initial copies failed");
        return fail(G4MI_E_HIP);
    }
    *out = h;
    return G4MI_OK;
}

int g4mi_destroy(g4mi_ctx *h) {
    if (!h) return G4MI_OK;
    Ctx *c = &h->c;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void *p : c->allocs) (void)hipFree(p);
    if (c->d_band) (void)hipFree(c->d_band);
    if (c->d_tilediag) (void)hipFree(c->d_tilediag);
    if (c->m_steps) (void)hipFree(c->m_steps);
    mix_clear_graphs(c);
    if (c->m_tune_i > 0 && !c->m_tuned)  // tuning started and not finished
        for (auto &e : c->m_tune_ev) (void)hipEventDestroy(e);
    for (auto &e : c->pool)
        if (e) (void)hipEventDestroy(e);
    if (c->incoh_ev) (void)hipEventDestroy(c->incoh_ev);
    if (c->h_recs) (void)hipHostFree(c->h_recs);
    if (c->h_bdg) (void)hipHostFree(c->h_bdg);
    if (c->m_rec_pin) (void)hipHostFree(c->m_rec_pin);
    for (auto &e : c->m_rec_ev)
        if (e) (void)hipEventDestroy(e);
    delete c->incoh;
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete h;
    return G4MI_OK;
}

int g4mi_set_incoherent(g4mi_ctx *h, int seed, int nbins, int do_loss, int do_spread) {
    if (!h || nbins < 1) {
        set_error("g4mi_set_incoherent: invalid argument");
        return G4MI_E_ARG;
    }
    Ctx *c = &h->c;
    if (c->precision != G4MI_FP64) {
        set_error("g4mi_set_incoherent: incoherent synchrotron radiation runs in the FP64 form only");
        return G4MI_E_ARG;
    }
    delete c->incoh;
    c->incoh = nullptr;
    c->incoh_loss = do_loss ? 1 : 0;
    c->incoh_spread = do_spread ? 1 : 0;
    if (!c->incoh_loss && !c->incoh_spread) return G4MI_OK;
    // Incoherent::init (:17-31) on rank 0: the stream is seeded with
    // round(1e9 u), u the first draw of ran2(seed)
    Ran2 rseed((unsigned int)seed);
    const double val = rseed.next() * 1e9;
    c->incoh = new Ran2((unsigned int)(int)std::round(val));
    // beamlet offsets: ceil(n_s / nbins) kicks per slice
    std::vector<long long> boff(c->ns + 1, 0);
    for (int s = 0; s < c->ns; s++) boff[s + 1] = boff[s] + (c->off[s + 1] - c->off[s] + nbins - 1) / nbins;
    if (c->d_bdg && (c->incoh_nbins != nbins)) {
        set_error("g4mi_set_incoherent: the beamlet size cannot change within a context");
        return G4MI_E_ARG;
    }
    c->incoh_nbins = nbins;
    c->incoh_nb = boff[c->ns];
    if (!c->d_bdg) {
        // capacity for any redistribution of the particles (one4one sorting)
        const size_t cap = (size_t)(c->ntot / nbins + c->ns + 1);
        G4_TRY(dalloc(c, &c->d_boff, (size_t)c->ns + 1));
        G4_TRY(dalloc(c, &c->d_bdg, cap));
        G4_HIP(hipHostMalloc((void **)&c->h_bdg, cap * sizeof(double)));
        G4_HIP(hipEventCreateWithFlags(&c->incoh_ev, hipEventDisableTiming));
        G4_HIP(hipEventRecord(c->incoh_ev, c->stream));
    }
    G4_HIP(hipMemcpy(c->d_boff, boff.data(), boff.size() * sizeof(long long), hipMemcpyHostToDevice));
    return G4MI_OK;
}

int g4mi_incoherent_state(g4mi_ctx *h, int *state, int store) {
    if (!h || !state) return G4MI_E_ARG;
    Ctx *c = &h->c;
    if (!c->incoh) {
        set_error("g4mi_incoherent_state: no incoherent synchrotron radiation stream in this context");
        return G4MI_E_ARG;
    }
    if (store) c->incoh->load(state);
    else c->incoh->save(state);
    return G4MI_OK;
}

int g4mi_add_harmonic(g4mi_ctx *hc, int harm, int ngrid, double gridmax, int *index) {
    if (!hc || harm < 2 || ngrid < 3 || !(gridmax > 0)) {
        set_error("g4mi_add_harmonic: invalid argument (harm >= 2, ngrid >= 3, gridmax > 0)");
        return G4MI_E_ARG;
    }
    Ctx *c = &hc->c;
    if (c->precision != G4MI_FP64) {
        set_error("g4mi_add_harmonic: harmonic fields run in the FP64 form only");
        return G4MI_E_ARG;
    }
    if (c->fft) {
        set_error("g4mi_add_harmonic: harmonic fields run with the ADI solver only");
        return G4MI_E_ARG;
    }
    if ((int)c->hf.size() >= kMaxHarm) {
        set_error("g4mi_add_harmonic: at most 4 harmonic fields besides the fundamental");
        return G4MI_E_ARG;
    }
    for (const Ctx::HField &e : c->hf)
        if (e.harm == harm) {
            set_error("g4mi_add_harmonic: a field at this harmonic exists already");
            return G4MI_E_ARG;
        }
    Ctx::HField h;
    h.harm = harm;
    h.ng = ngrid;
    h.gridmax = gridmax;
    h.dgrid = 2 * gridmax / (double)(ngrid - 1);  // Field::init :45-46
    h.xks = 4. * asin(1) / (c->lambda0 / (double)harm);
    const size_t cells = (size_t)c->ns * ngrid * ngrid;
    G4_TRY(dalloc(c, &h.d_field, cells));
    G4_TRY(dalloc(c, &h.d_work, cells));
    G4_TRY(dalloc(c, &h.d_src, cells));
    G4_TRY(dalloc(c, &h.d_fac, 3 * (size_t)ngrid));
    G4_TRY(dalloc(c, &h.d_box, (size_t)c->ns));
    G4_HIP(hipMemsetAsync(h.d_field, 0, cells * sizeof(double2), c->stream));
    G4_HIP(hipStreamSynchronize(c->stream));
    c->hf.push_back(h);
    if (index) *index = (int)c->hf.size() - 1;
    return G4MI_OK;
}

static int harm_of(g4mi_ctx *hc, int index, Ctx::HField **h) {
    if (!hc || index < 0 || index >= (int)hc->c.hf.size()) {
        set_error("harmonic field index out of range");
        return G4MI_E_ARG;
    }
    *h = &hc->c.hf[index];
    return G4MI_OK;
}

int g4mi_harmonic_upload(g4mi_ctx *hc, int index, const double *field) {
    Ctx::HField *h = nullptr;
    G4_TRY(harm_of(hc, index, &h));
    if (!field) return G4MI_E_ARG;
    G4_HIP(hipMemcpy(h->d_field, field, (size_t)hc->c.ns * h->ng * h->ng * sizeof(double2), hipMemcpyHostToDevice));
    return G4MI_OK;
}

int g4mi_harmonic_download(g4mi_ctx *hc, int index, double *field) {
    Ctx::HField *h = nullptr;
    G4_TRY(harm_of(hc, index, &h));
    if (!field) return G4MI_E_ARG;
    G4_HIP(hipStreamSynchronize(hc->c.stream));
    G4_HIP(hipMemcpy(field, h->d_field, (size_t)hc->c.ns * h->ng * h->ng * sizeof(double2), hipMemcpyDeviceToHost));
    return G4MI_OK;
}

int g4mi_diag_bunching(g4mi_ctx *hc, int hmax, double *bunching, double *phase) {
    if (!hc || hmax < 2 || !bunching || !phase) {
        set_error("g4mi_diag_bunching: invalid argument (hmax >= 2, outputs)");
        return G4MI_E_ARG;
    }
    Ctx *c = &hc->c;
    if (c->precision != G4MI_FP64) {
        set_error("g4mi_diag_bunching: harmonic bunching runs in the FP64 form only");
        return G4MI_E_ARG;
    }
    const int ns = c->ns;
    if (!c->d_bunch) G4_TRY(dalloc(c, &c->d_bunch, (size_t)ns * 16));
    std::vector<double> hb((size_t)ns * 16);
    for (int h0 = 2; h0 <= hmax; h0 += 8) {
        const int cnt = std::min(8, hmax - h0 + 1);
        launch_diag_bunch(c->stream, ns, c->d_off, c->d_th, h0, cnt, c->d_bunch);
        G4_HIP(hipGetLastError());
        G4_HIP(hipMemcpyAsync(hb.data(), c->d_bunch, (size_t)ns * cnt * 2 * sizeof(double), hipMemcpyDeviceToHost,
                              c->stream));
        G4_HIP(hipStreamSynchronize(c->stream));
        for (int is = 0; is < ns; is++) {
            const long long n = c->off[is + 1] - c->off[is];
            const double norm = n > 0 ? 1. / (double)n : 1.;  // DiagBeam: 1 / slice size
            for (int k = 0; k < cnt; k++) {
                const std::complex<double> b(hb[2 * ((size_t)is * cnt + k)] * norm,
                                             hb[2 * ((size_t)is * cnt + k) + 1] * norm);
                const size_t o = (size_t)(h0 + k - 2) * ns + is;
                bunching[o] = std::abs(b);
                phase[o] = atan2(b.imag(), b.real());
            }
        }
    }
    return G4MI_OK;
}

int g4mi_diag_harmonic(g4mi_ctx *hc, int index, g4mi_diag_out *out) {
    Ctx::HField *h = nullptr;
    G4_TRY(harm_of(hc, index, &h));
    if (!out) return G4MI_E_ARG;
    Ctx *c = &hc->c;
    const int ns = c->ns;
    launch_diag_field(c->stream, ns, h->d_field, h->ng, c->d_diag_f);
    std::vector<double> hf(8 * (size_t)ns);
    G4_HIP(hipMemcpyAsync(hf.data(), c->d_diag_f, hf.size() * 8, hipMemcpyDeviceToHost, c->stream));
    G4_HIP(hipStreamSynchronize(c->stream));
    c->field_diag_fresh = false;  // the fundamental's sums in d_diag_f were overwritten
    finish_field(hf, ns, c->first, h->dgrid, h->xks, c->slicelength, out);
    return G4MI_OK;
}

int g4mi_harmonic_power(g4mi_ctx *hc, int index, double *power) {
    g4mi_diag_out o{};
    o.power = power;
    return g4mi_diag_harmonic(hc, index, &o);
}

int g4mi_set_sorting(g4mi_ctx *h, int one4one) {
    if (!h) {
        set_error("g4mi_set_sorting: null context");
        return G4MI_E_ARG;
    }
    Ctx *c = &h->c;
    if (one4one && c->precision != G4MI_FP64) {
        set_error("g4mi_set_sorting: one4one beams run in the FP64 form only");
        return G4MI_E_ARG;
    }
    c->one4one = one4one != 0;
    return G4MI_OK;
}

int g4mi_sort(g4mi_ctx *h, long long *nkept) {
    if (!h) {
        set_error("g4mi_sort: null context");
        return G4MI_E_ARG;
    }
    Ctx *c = &h->c;
    if (!c->one4one) {  // Beam::sort does nothing for beams that are not one4one
        if (nkept) *nkept = c->ntot;
        return G4MI_OK;
    }
    G4_TRY(sort_beam(c));
    if (nkept) *nkept = c->ntot;
    return G4MI_OK;
}

int g4mi_get_slices(g4mi_ctx *h, long long *off, double *current) {
    if (!h) {
        set_error("g4mi_get_slices: null context");
        return G4MI_E_ARG;
    }
    Ctx *c = &h->c;
    if (off) std::copy(c->off.begin(), c->off.end(), off);
    if (current) {
        if ((int)c->h_cur.size() != c->ns) {
            set_error("g4mi_get_slices: no beam uploaded");
            return G4MI_E_STATE;
        }
        std::copy(c->h_cur.begin(), c->h_cur.end(), current);
    }
    return G4MI_OK;
}

int g4mi_set_wake(g4mi_ctx *h, const double *eloss) {
    if (!h) {
        set_error("g4mi_set_wake: null context");
        return G4MI_E_ARG;
    }
    Ctx *c = &h->c;
    if (c->precision != G4MI_FP64) {
        set_error("g4mi_set_wake: wakes run in the FP64 form only");
        return G4MI_E_ARG;
    }
    if (!eloss) {
        c->d_wake = nullptr;  // buffers stay allocated for a later call
        c->h_wake.clear();
        return G4MI_OK;
    }
    if (!c->d_wake_dg) {
        G4_TRY(dalloc(c, &c->d_wake_dg, (size_t)c->ns));
        G4_TRY(dalloc(c, &c->d_wake_buf, (size_t)c->ns));
    }
    c->h_wake.assign(eloss, eloss + c->ns);
    G4_HIP(hipMemcpy(c->d_wake_buf, eloss, c->ns * sizeof(double), hipMemcpyHostToDevice));
    c->d_wake = c->d_wake_buf;
    c->wake_dz = std::nan("");
    return G4MI_OK;
}

int g4mi_set_efield(g4mi_ctx *h, int longrange, int nz, int nphi, int ngrid_r, double rmax) {
    if (!h || nz < 0 || nphi < 0 || ngrid_r < 1 || ngrid_r > 256) {
        set_error("g4mi_set_efield: invalid argument (radial grid limited to 256)");
        return G4MI_E_ARG;
    }
    Ctx *c = &h->c;
    c->longrange = longrange ? 1 : 0;
    c->sc_nz = nz;
    c->sc_nphi = nphi;
    c->sc_ng = ngrid_r;
    c->sc_rmax = rmax;
    G4_HIP(hipMemcpyAsync(c->d_rmax, &c->sc_rmax, sizeof(double), hipMemcpyHostToDevice, c->stream));
    G4_HIP(hipStreamSynchronize(c->stream));
    if (!c->longrange) {
        G4_HIP(hipMemsetAsync(c->d_esc, 0, c->ns * sizeof(double), c->stream));
        G4_HIP(hipStreamSynchronize(c->stream));
    }
    return G4MI_OK;
}

int g4mi_set_field_solver(g4mi_ctx *h, int fft, int filter, double xcut, double ycut, double sigmoid) {
    if (!h) {
        set_error("g4mi_set_field_solver: null context");
        return G4MI_E_ARG;
    }
    Ctx *c = &h->c;
    if (fft && c->precision != G4MI_FP64) {
        set_error("g4mi_set_field_solver: the spectral solver runs in the FP64 form only");
        return G4MI_E_ARG;
    }
    if (fft && !c->hf.empty()) {
        set_error("g4mi_set_field_solver: the spectral solver does not run harmonic fields");
        return G4MI_E_ARG;
    }
    c->fft = fft ? 1 : 0;
    // FieldSolverFFT::initSourceFilter (:159-170): no filter for sig, xc or yc <= 0
    c->fft_filter = (fft && filter && sigmoid > 0 && xcut > 0 && ycut > 0) ? 1 : 0;
    c->fft_xc = xcut;
    c->fft_yc = ycut;
    c->fft_sig = sigmoid;
    c->fft_ready = false;
    c->field_diag_fresh = false;
    if (c->fft_filter) {
        const size_t nn = (size_t)c->ng * c->ng;
        G4_TRY(ensure_dft(c));
        if (!c->d_fft_sig) G4_TRY(dalloc(c, &c->d_fft_sig, nn));
        std::vector<double> sg(nn);
        fft_sigmoid(c->ng, xcut, ycut, sigmoid, sg.data());
        G4_HIP(hipMemcpy(c->d_fft_sig, sg.data(), sg.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    return G4MI_OK;
}

int g4mi_beam_upload(g4mi_ctx *h, const double *g, const double *th, const double *x,
                     const double *y, const double *px, const double *py, const double *cur) {
    if (!h || !g || !th || !x || !y || !px || !py || !cur) {
        set_error("g4mi_beam_upload: null pointer");
        return G4MI_E_ARG;
    }
    Ctx *c = &h->c;
    if (c->precision == G4MI_FP32) return mix_beam_upload(c, g, th, x, y, px, py, cur);
    const size_t nb = (size_t)c->ntot * sizeof(double);
    const double *src[6] = {g, th, x, y, px, py};
    double *dst[6] = {c->d_g, c->d_th, c->d_x, c->d_y, c->d_px, c->d_py};
    for (int k = 0; k < 6; k++)
        if (nb) G4_HIP(hipMemcpyAsync(dst[k], src[k], nb, hipMemcpyHostToDevice, c->stream));
    G4_HIP(hipMemcpyAsync(c->d_cur, cur, c->ns * sizeof(double), hipMemcpyHostToDevice, c->stream));
    G4_HIP(hipStreamSynchronize(c->stream));
    c->h_cur.assign(cur, cur + c->ns);
    c->has_beam = true;
    return G4MI_OK;
}

int g4mi_load_beam_device(g4mi_ctx *h, const g4mi_slice_beam *slices, int nbins, int shotnoise, int seed,
                          int islice0, double lambda0, double sample, int timerun) {
    if (!h || !slices || nbins < 1) {
        set_error("g4mi_load_beam_device: invalid argument");
        return G4MI_E_ARG;
    }
    Ctx *c = &h->c;
    const int ns = c->ns;
    const long long n = c->ntot;
    int maxpart = 0;
    for (int j = 0; j < ns; j++) {
        const long long m = c->off[j + 1] - c->off[j];
        if (m % nbins != 0) {
            set_error("g4mi_load_beam_device: NPART is not a multiple of NBINS");
            return G4MI_E_ARG;
        }
        maxpart = std::max(maxpart, (int)m);
    }
    // host: the per-slice sequence starts and the sequential shot-noise
    // stream (LoadBeam.cpp:127-212), in the order g4mi_load_beam draws them
    std::vector<unsigned int> starts(ns);
    for (int j = 0; j < ns; j++) starts[j] = quiet_start(j + islice0);
    const int nh = (nbins - 1) / 2;
    std::vector<double> draws;
    std::vector<long long> dbase(ns, -1);
    if (shotnoise && timerun && nh > 0) {
        draws.reserve(2 * (size_t)nh * (size_t)(n / nbins));
        Ran2 stream(shot_noise_seed(seed, 0));
        for (int j = 0; j < ns; j++) {
            const long long before = (long long)draws.size() / 2;
            const double ne = slices[j].current * lambda0 * sample / kCeH();
            if (shot_noise_draws(stream, (int)(c->off[j + 1] - c->off[j]), nbins, ne, draws) > 0)
                dbase[j] = before;
        }
    }
    // one temporary block: starts, slices, dbase, draws, and in the mixed
    // form the FP64 coordinates and the energy partial sums
    const bool mixed = c->precision == G4MI_FP32;
    constexpr int kSumBlocks = 1024;
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t o_slices = up(ns * sizeof(unsigned int));
    const size_t o_dbase = o_slices + up(ns * sizeof(g4mi_slice_beam));
    const size_t o_draws = o_dbase + up(ns * sizeof(long long));
    const size_t o_coord = o_draws + up(std::max<size_t>(draws.size(), 1) * sizeof(double));
    const size_t o_sums = o_coord + (mixed ? up(6 * (size_t)n * sizeof(double)) : 0);
    const size_t total = o_sums + (mixed ? up(kSumBlocks * sizeof(double)) : 0);
    char *tmp = nullptr;
    G4_HIP(hipMalloc(&tmp, total));
    struct Free {
        char *p;
        ~Free() { (void)hipFree(p); }
    } free_tmp{tmp};
    unsigned int *d_starts = (unsigned int *)tmp;
    g4mi_slice_beam *d_slices = (g4mi_slice_beam *)(tmp + o_slices);
    long long *d_dbase = (long long *)(tmp + o_dbase);
    double *d_draws = (double *)(tmp + o_draws);
    G4_HIP(hipMemcpyAsync(d_starts, starts.data(), ns * sizeof(unsigned int), hipMemcpyHostToDevice, c->stream));
    G4_HIP(hipMemcpyAsync(d_slices, slices, ns * sizeof(g4mi_slice_beam), hipMemcpyHostToDevice, c->stream));
    G4_HIP(hipMemcpyAsync(d_dbase, dbase.data(), ns * sizeof(long long), hipMemcpyHostToDevice, c->stream));
    if (!draws.empty())
        G4_HIP(hipMemcpyAsync(d_draws, draws.data(), draws.size() * sizeof(double), hipMemcpyHostToDevice,
                              c->stream));
    double *g = c->d_g, *th = c->d_th, *x = c->d_x, *y = c->d_y, *px = c->d_px, *py = c->d_py;
    if (mixed) {
        g = (double *)(tmp + o_coord);
        th = g + n;
        x = th + n;
        y = x + n;
        px = y + n;
        py = px + n;
    }
    const double theta0 = 4. * asin(1.);  // LoadBeam.cpp:127
    launch_quiet_load(c->stream, ns, c->d_off, d_starts, d_slices, nbins, theta0, g, th, x, y, px, py);
    G4_HIP(hipGetLastError());
    if (!draws.empty()) {
        launch_shot_noise(c->stream, ns, c->d_off, d_dbase, nbins, nh, d_draws, th, maxpart);
        G4_HIP(hipGetLastError());
    }
    std::vector<double> cur(ns);
    for (int j = 0; j < ns; j++) cur[j] = slices[j].current;
    G4_HIP(hipMemcpyAsync(c->d_cur, cur.data(), ns * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (mixed) {
        // offset form as mix_beam_upload: gref = the mean energy rounded to
        // FP32, a phase base per slice
        double *d_sums = (double *)(tmp + o_sums);
        std::vector<double> sums(kSumBlocks);
        launch_sum_gamma(c->stream, g, n, d_sums, kSumBlocks);
        G4_HIP(hipGetLastError());
        G4_HIP(hipMemcpyAsync(sums.data(), d_sums, kSumBlocks * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        G4_HIP(hipStreamSynchronize(c->stream));
        double gsum = 0;
        for (double v : sums) gsum += v;
        c->gref = n > 0 ? (double)(float)(gsum / (double)n) : 1.0;
        launch_theta_base(c->stream, ns, c->d_off, th, c->m_base);
        G4_HIP(hipGetLastError());
        launch_to_offsets(c->stream, ns, c->d_off, n, c->m_base, c->gref, g, th, x, y, px, py, c->m_dg, c->m_dth,
                          c->m_x, c->m_y, c->m_px, c->m_py, maxpart);
        G4_HIP(hipGetLastError());
        mix_clear_graphs(c);  // gref is part of every captured launch
    }
    G4_HIP(hipStreamSynchronize(c->stream));
    c->h_cur = cur;
    c->has_beam = true;
    return G4MI_OK;
}

int g4mi_beam_download(g4mi_ctx *h, double *g, double *th, double *x, double *y, double *px,
                       double *py) {
    if (!h) return G4MI_E_ARG;
    Ctx *c = &h->c;
    if (c->precision == G4MI_FP32) return mix_beam_download(c, g, th, x, y, px, py);
    const size_t nb = (size_t)c->ntot * sizeof(double);
    double *dst[6] = {g, th, x, y, px, py};
    const double *src[6] = {c->d_g, c->d_th, c->d_x, c->d_y, c->d_px, c->d_py};
    for (int k = 0; k < 6; k++)
        if (dst[k] && nb) G4_HIP(hipMemcpyAsync(dst[k], src[k], nb, hipMemcpyDeviceToHost, c->stream));
    G4_HIP(hipStreamSynchronize(c->stream));
    return G4MI_OK;
}

int g4mi_field_upload(g4mi_ctx *h, const double *field, int first) {
    if (!h || !field) return G4MI_E_ARG;
    Ctx *c = &h->c;
    if (first < 0 || first >= c->ns) {
        set_error("g4mi_field_upload: first out of range");
        return G4MI_E_ARG;
    }
    if (c->precision == G4MI_FP32) return mix_field_upload(c, field, first);
    G4_HIP(hipMemcpyAsync(c->d_field, field, (size_t)c->ns * c->ng * c->ng * sizeof(double2),
                          hipMemcpyHostToDevice, c->stream));
    G4_HIP(hipStreamSynchronize(c->stream));
    c->first = first;
    c->has_field = true;
    c->field_diag_fresh = false;
    return G4MI_OK;
}

int g4mi_field_download(g4mi_ctx *h, double *field, int *first) {
    if (!h) return G4MI_E_ARG;
    Ctx *c = &h->c;
    if (c->precision == G4MI_FP32) return mix_field_download(c, field, first);
    if (field)
        G4_HIP(hipMemcpyAsync(field, c->d_field, (size_t)c->ns * c->ng * c->ng * sizeof(double2),
                              hipMemcpyDeviceToHost, c->stream));
    G4_HIP(hipStreamSynchronize(c->stream));
    if (first) *first = c->first;
    return G4MI_OK;
}

int g4mi_beam_track(g4mi_ctx *h, const g4mi_step_params *st, double gammaref, double delz) {
    if (!h || !st) return G4MI_E_ARG;
    if (h->c.precision == G4MI_FP32) return mix_split_call("g4mi_beam_track");
    return beam_track(&h->c, *st, gammaref, delz);
}

int g4mi_field_track(g4mi_ctx *h, const g4mi_step_params *st, double delz) {
    if (!h || !st) return G4MI_E_ARG;
    if (h->c.precision == G4MI_FP32) return mix_split_call("g4mi_field_track");
    return field_track(&h->c, *st, delz);
}

int g4mi_slippage(g4mi_ctx *h, double slip) {
    if (!h) return G4MI_E_ARG;
    if (h->c.precision == G4MI_FP32) return mix_split_call("g4mi_slippage");
    return slippage(&h->c, slip);
}

int g4mi_step(g4mi_ctx *h, const g4mi_step_params *st, double gammaref) {
    if (!h || !st) return G4MI_E_ARG;
    if (h->c.precision == G4MI_FP32) {
        int done = 0;
        return mix_run(&h->c, st, 1, gammaref, nullptr, &done);
    }
    G4_TRY(step(&h->c, *st, gammaref));
    return step_end_timing(&h->c);
}

// Deferred records of the FP64 run loop: each record's beam and field sums
// (and, when asked for, the auxiliary sums with the space-charge fields and
// the spectral moments) are written into a pinned slot by copy kernels on
// the step's stream, and the host finishes them when the slots are full or
// the run ends, so a record after every step does not stall the device
// between steps.  one4one runs (whose slices change at sorts) take the
// synchronous path.
static bool record_deferrable(const Ctx *c, const g4mi_diag_out *o) {
    (void)o;
    return !c->one4one;
}

static int flush_records(Ctx *c) {
    if (c->pend.empty()) return G4MI_OK;
    G4_HIP(hipStreamSynchronize(c->stream));
    const size_t ns = (size_t)c->ns, row = kRecRow * ns;
    std::vector<double> hb(8 * ns), hf(8 * ns);
    for (size_t k = 0; k < c->pend.size(); k++) {
        Ctx::PendingRecord &r = c->pend[k];
        const double *slot = c->h_recs + k * row;
        std::copy(slot, slot + 8 * ns, hb.begin());
        std::copy(slot + 8 * ns, slot + 16 * ns, hf.begin());
        if (r.fresh)  // slices zeroed by slippage after the ADI's sums
            for (int z : r.zeroed)
                for (int q = 0; q < 8; q++) hf[8 * (size_t)z + q] = 0;
        finish_diag(c, hb, hf, r.first, &r.out);
        if (wants_aux(&r.out)) finish_aux(c, slot + 16 * ns, slot + 32 * ns, slot + 33 * ns, r.wake, &r.out);
        if (wants_spectral(&r.out)) finish_spectral(c, slot + 34 * ns, r.first, &r.out);
    }
    c->pend.clear();
    return G4MI_OK;
}

static int enqueue_record(Ctx *c, const g4mi_diag_out *o) {
    const size_t ns = (size_t)c->ns, row = kRecRow * ns;
    if (!c->h_recs) G4_TRY(alloc_records(c));
    if ((int)c->pend.size() == c->rec_cap) G4_TRY(flush_records(c));
    // the beam sums go straight into the slot; the field sums of the ADI
    // (or of k_diag_field) are copied after them
    double *slot = c->d_recs + c->pend.size() * row;
    if (c->beamsum_fresh)  // the fused main kernel summed the step's particles
        launch_copy_out(c->stream, c->d_beamsum, slot, (long long)(8 * ns));
    else
        launch_diag_beam(c->stream, c->ns, c->d_off, c->d_g, c->d_th, c->d_x, c->d_y, slot);
    if (!c->field_diag_fresh) launch_diag_field(c->stream, c->ns, c->d_field, c->ng, c->d_diag_f);
    launch_copy_out(c->stream, c->d_diag_f, slot + 8 * ns, (long long)(8 * ns));
    std::vector<double> wake;
    if (wants_aux(o)) {
        G4_TRY(launch_aux(c));
        launch_copy_out(c->stream, c->d_aux, slot + 16 * ns, (long long)(16 * ns));
        launch_copy_out(c->stream, c->d_esc, slot + 32 * ns, (long long)ns);
        launch_copy_out(c->stream, c->d_efield, slot + 33 * ns, (long long)ns);
        wake = c->wake_applied;
    }
    if (wants_spectral(o)) {
        G4_TRY(spectral_moments(c, c->d_field));
        launch_copy_out(c->stream, c->d_fft_mom, slot + 34 * ns, (long long)(5 * ns));
    }
    G4_HIP(hipGetLastError());
    c->pend.push_back({*o, c->first, c->field_diag_fresh, c->zeroed_since_adi, std::move(wake)});
    return G4MI_OK;
}

int g4mi_run(g4mi_ctx *h, const g4mi_step_params *steps, int nsteps, double gammaref,
             g4mi_diag_out *records) {
    if (!h || (!steps && nsteps > 0) || nsteps < 0) return G4MI_E_ARG;
    int done = 0;
    if (h->c.precision == G4MI_FP32) {
        G4_TRY(mix_run(&h->c, steps, nsteps, gammaref, records, &done));
        return done;
    }
    Ctx *c = &h->c;
    int rc = G4MI_OK;
    for (int i = 0; i < nsteps && rc == G4MI_OK; i++) {
        if (steps[i].marker & 8) break;
        rc = step(c, steps[i], gammaref);
        if (rc == G4MI_OK && records) {
            if (record_deferrable(c, &records[i])) {
                rc = enqueue_record(c, &records[i]);
            } else {
                rc = flush_records(c);
                if (rc == G4MI_OK) rc = diag(c, &records[i]);
            }
        }
        if (rc == G4MI_OK) rc = step_end_timing(c);
        if (rc == G4MI_OK) done++;
    }
    const int rf = flush_records(c);
    if (rc != G4MI_OK) return rc;
    if (rf != G4MI_OK) return rf;
    G4_HIP(hipStreamSynchronize(c->stream));
    return done;
}

int g4mi_diag(g4mi_ctx *h, g4mi_diag_out *out) {
    if (!h || !out) return G4MI_E_ARG;
    if (h->c.precision == G4MI_FP32) return mix_diag(&h->c, out);
    return diag(&h->c, out);
}

int g4mi_sort_index(g4mi_ctx *h, double slen, int *atar, long long *moved) {
    if (!h || slen == 0) return G4MI_E_ARG;
    Ctx *c = &h->c;
    if (!c->d_atar) G4_TRY(dalloc(c, &c->d_atar, (size_t)c->ntot));
    G4_HIP(hipMemsetAsync(c->d_count, 0, sizeof(unsigned long long), c->stream));
    if (c->precision == G4MI_FP32)
        launch_mix_sort_index(c->stream, mix_arrays(c), 1. / slen, c->d_atar, c->d_count);
    else
        launch_sort_index(c->stream, c->ntot, c->d_th, 1. / slen, c->d_atar, c->d_count);
    unsigned long long m = 0;
    if (atar && c->ntot)
        G4_HIP(hipMemcpyAsync(atar, c->d_atar, c->ntot * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    G4_HIP(hipMemcpyAsync(&m, c->d_count, sizeof(m), hipMemcpyDeviceToHost, c->stream));
    G4_HIP(hipStreamSynchronize(c->stream));
    if (moved) *moved = (long long)m;
    return G4MI_OK;
}

// g4mi_sort_index without any allocation: the target indices and the
// counter live in the library's static device scratch (see `scratch`).
// Calls that overlap on one GPU share that scratch and race.
int g4mi_sort_index_noalloc(g4mi_ctx *h, double slen, int *atar, long long *moved) {
    if (!h || slen == 0) return G4MI_E_ARG;
    Ctx *c = &h->c;
    if (c->ntot > kNoAllocMaxParticles) {
        set_error("g4mi_sort_index_noalloc: " + std::to_string(c->ntot) +
                  " particles exceed the static scratch's maximum of " +
                  std::to_string(kNoAllocMaxParticles) + "; use g4mi_sort_index");
        return G4MI_E_ARG;
    }
    if (c->device < 0 || c->device >= kMaxDevices || !g_scratch_ptr[c->device]) {
        set_error(std::string("g4mi_sort_index_noalloc: no scratch address on this device: ") +
                  hipGetErrorString(c->device >= 0 && c->device < kMaxDevices ? g_scratch_err[c->device]
                                                                               : hipErrorInvalidDevice));
        return G4MI_E_HIP;
    }
    char *base = (char *)g_scratch_ptr[c->device];
    auto *cnt = (unsigned long long *)base;
    int *d_at = (int *)(base + kScratchAtar);
    // the counter starts at zero: a fill kernel, not a memset node, so the
    // sequence stays valid under stream capture
    hipLaunchKernelGGL(k_clear_u64, dim3(1), dim3(64), 0, c->stream, cnt);
    if (c->precision == G4MI_FP32)
        launch_mix_sort_index(c->stream, mix_arrays(c), 1. / slen, d_at, cnt);
    else
        launch_sort_index(c->stream, c->ntot, c->d_th, 1. / slen, d_at, cnt);
    G4_HIP(hipGetLastError());
    unsigned long long m = 0;
    if (atar && c->ntot)
        G4_HIP(hipMemcpyAsync(atar, d_at, (size_t)c->ntot * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    G4_HIP(hipMemcpyAsync(&m, cnt, sizeof(m), hipMemcpyDeviceToHost, c->stream));
    G4_HIP(hipStreamSynchronize(c->stream));
    if (moved) *moved = (long long)m;
    return G4MI_OK;
}

int g4mi_set_schedule(g4mi_ctx *h, const char *name, int value) {
    if (!h || !name) return G4MI_E_ARG;
    Ctx *c = &h->c;
    if (std::string(name) == "adi_kernel" && value >= -1 && value <= 4) {
        c->adi_kernel = value;
        c->adi_fixed = value >= 0;
        if (value < 0) c->adi_choice.clear();
        return G4MI_OK;
    }
    if (std::string(name) == "fft_variant" && value >= -1 && value < kCgemmVariants) {
        c->fft_variant = value;  // -1: tune again at the next spectral step
        return G4MI_OK;
    }
    if (std::string(name) == "fft_propagation" && value >= -1 && value <= 1) {
        if (value == 1 && !fft_conv_length(c->ng)) {
            set_error("g4mi_set_schedule: fft_propagation 1 supports ngrid <= 256");
            return G4MI_E_ARG;
        }
        c->fft_prop = value;  // -1: tune again at the next spectral step
        return G4MI_OK;
    }
    if (std::string(name) == "adi_refine" && (value == 0 || value == 1)) {
        c->m_refine = value != 0;
        return G4MI_OK;
    }
    if (std::string(name) == "mixed_adi" && value >= kAdiTile && value <= kAdiWide) {
        c->m_adi_mode = value;
        c->m_autotune = false;
        return G4MI_OK;
    }
    if (std::string(name) == "main_nt" && (value == 256 || value == 384 || value == 512)) {
        c->m_main_nt = value;
        c->m_autotune = false;
        return G4MI_OK;
    }
    if (std::string(name) == "main_lds" && value >= 0 && value <= 2) {
        c->m_main_lds = value;
        c->m_autotune = false;
        return G4MI_OK;
    }
    if (std::string(name) == "fp64_fused" && value >= -1 && value <= 1) {
        c->fuse64 = value;
        return G4MI_OK;
    }
    if (std::string(name) == "persistent" && (value == 0 || value == 1)) {
        c->m_persistent = value != 0;
        if (value && c->m_adi_mode == kAdiTile) c->m_adi_mode = kAdiWide;  // the persistent step streams strips
        c->m_autotune = false;
        return G4MI_OK;
    }
    if (std::string(name) == "autotune" && (value == 0 || value == 1)) {
        c->m_autotune = value != 0;
        if (c->m_tune_i > 0 && !c->m_tuned)
            for (auto &e : c->m_tune_ev) (void)hipEventDestroy(e);
        c->m_tuned = false;
        c->m_tune_i = 0;
        for (int q = 0; q < kMainCands; q++) c->m_tune_main[q] = 0, c->m_tune_nmain[q] = 0;
        for (int q = 0; q < 3; q++) c->m_tune_adi[q] = 0, c->m_tune_nadi[q] = 0;
        c->m_tune_graph = -1;
        return G4MI_OK;
    }
    if (std::string(name) == "graphs" && (value == 0 || value == 1)) {
        c->m_graphs = value != 0;
        return G4MI_OK;
    }
    if (std::string(name) == "mixed_fuse_front" && (value == 0 || value == 1)) {
        c->m_fuse_front = value != 0;
        return G4MI_OK;
    }
    set_error(std::string("g4mi_set_schedule: unknown entry or value: ") + name);
    return G4MI_E_ARG;
}

int g4mi_get_schedule(g4mi_ctx *h, char *buf, int len) {
    if (!h || !buf || len < 1) return G4MI_E_ARG;
    Ctx *c = &h->c;
    char tmp[4096];
    static const char *names[5] = {"thomas", "fused", "fused-prefetch", "band", "band-wide"};
    if (c->precision == G4MI_FP32) {
        std::string bands = "[";
        for (size_t k = 0; k < c->h_binfo.size(); k++)
            bands += (k ? ", " : "") + std::to_string(c->h_binfo[k].m) + "/" + std::to_string(c->band_tmpl[k]);
        bands += "]";
        snprintf(tmp, sizeof(tmp),
                 "{\"precision\": \"fp32\", \"adi\": \"%s\", \"adi_refine\": %d, "
                 "\"main_nt\": %d, \"main_lds_kib\": %d, \"persistent\": %d, \"autotune\": %d, \"tuned\": %d, "
                 "\"graphs\": %d, \"fuse_front\": %d, \"bands_m_over_template\": \"%s\", \"graphs_captured\": %d, "
                 "\"tuning\": %s, \"fno_skipped_steps\": %lld}",
                 c->m_adi_mode == kAdiWide ? "band32-wide" : c->m_adi_mode == kAdiStrips ? "band32-stream" : "band32-tile",
                 c->m_refine ? 1 : 0, c->m_main_nt,
                 (c->m_main_lds == 2 && c->m_main_nt != 512) ? 32 : c->m_main_lds ? 44 : 68, c->m_persistent ? 1 : 0, c->m_autotune ? 1 : 0, c->m_tuned ? 1 : 0, c->m_graphs ? 1 : 0,
                 c->m_fuse_front ? 1 : 0, bands.c_str(), (int)c->m_graph.size(), c->m_tune_log.empty() ? "null" : c->m_tune_log.c_str(),
                 c->fno_skipped);
        snprintf(buf, (size_t)len, "%s", tmp);
        return G4MI_OK;
    }
    if (c->fft) {
        snprintf(tmp, sizeof(tmp),
                 "{\"field_solver\": \"fft\", \"source_filter\": %d, \"gemm_tile\": %d, \"gemm_variant\": %d, "
                 "\"gemm_tuned_ms\": [%.4f, %.4f, %.4f, %.4f], \"propagation\": \"%s\", "
                 "\"conv_length\": %d, \"conv_tuned_ms\": %.4f, \"fp64_fused\": %d}",
                 c->fft_filter, cgemm_tile(c->ng), c->fft_variant, c->fft_ms[0], c->fft_ms[1], c->fft_ms[2],
                 c->fft_ms[3], c->fft_prop == 1 ? "fft-conv" : (c->fft_prop == 0 ? "gemm" : "untuned"),
                 fft_conv_length(c->ng), c->fft_conv_ms, fp64_fuse_on(c) ? 1 : 0);
        snprintf(buf, (size_t)len, "%s", tmp);
        return G4MI_OK;
    }
    std::string bydz = "{";
    for (const auto &kv : c->adi_choice) {
        char e[64];
        snprintf(e, sizeof(e), "%s\"%.6g\": \"%s\"", bydz.size() > 1 ? ", " : "", kv.first, names[kv.second]);
        bydz += e;
    }
    bydz += "}";
    snprintf(tmp, sizeof(tmp),
             "{\"field_solver\": \"adi\", \"adi_kernel\": %d, \"adi\": \"%s\", \"adi_window\": %d, "
             "\"adi_band_m\": %d, \"adi_tuned_ms\": [%.4f, %.4f, %.4f, %.4f, %.4f], \"adi_by_delz\": %s, "
             "\"fp64_fused\": %d}",
             c->adi_kernel, c->adi_kernel >= 0 ? names[c->adi_kernel] : "untuned", c->adi_tile,
             c->adi_band_m, c->adi_ms[0], c->adi_ms[1], c->adi_ms[2], c->adi_ms[3], c->adi_ms[4], bydz.c_str(),
             fp64_fuse_on(c) ? 1 : 0);
    snprintf(buf, (size_t)len, "%s", tmp);
    return G4MI_OK;
}

int g4mi_attach_fno(g4mi_ctx *h, g4mi_fno *fh, int mode, int every, double threshold) {
    if (!h || mode < 0 || mode > 2 || (mode && (!fh || every < 1 || !(threshold >= 0)))) {
        set_error("g4mi_attach_fno: invalid argument");
        return G4MI_E_ARG;
    }
    Ctx *c = &h->c;
    if (mode && c->precision != G4MI_FP32) {
        set_error("g4mi_attach_fno: the learned field step needs a mixed-form (G4MI_FP32) context");
        return G4MI_E_STATE;
    }
    if (mode && fno_ngrid(fno_of(fh)) != c->ng) {
        set_error("g4mi_attach_fno: network grid differs from the context's");
        return G4MI_E_ARG;
    }
    c->fno = mode ? fno_of(fh) : nullptr;
    c->fno_mode = mode;
    c->fno_every = every;
    c->fno_thr = threshold;
    c->fno_steps = c->fno_adi_steps = c->fno_nres_adi = c->fno_njudge = c->fno_nres_all = 0;
    c->fno_reject_run = c->fno_skip = 0;
    c->fno_skipped = 0;
    c->fno_res_sum = c->fno_res_adi_sum = c->fno_judge_field = c->fno_judge_update = c->fno_res_all_sum = 0;
    return G4MI_OK;
}

int g4mi_fno_judge(g4mi_ctx *h, long long *njudge, double *rel_field_err, double *rel_update_err,
                   double *mean_residual_all) {
    if (!h) return G4MI_E_ARG;
    const Ctx *c = &h->c;
    const double n = (double)std::max(1LL, c->fno_njudge);
    if (njudge) *njudge = c->fno_njudge;
    if (rel_field_err) *rel_field_err = c->fno_judge_field / n;
    if (rel_update_err) *rel_update_err = c->fno_judge_update / n;
    if (mean_residual_all)
        *mean_residual_all = c->fno_nres_all ? c->fno_res_all_sum / (double)c->fno_nres_all : 0.0;
    return G4MI_OK;
}

int g4mi_fno_stats(g4mi_ctx *h, long long *fno_steps, long long *adi_steps, double *mean_res_fno,
                   double *mean_res_adi) {
    if (!h) return G4MI_E_ARG;
    Ctx *c = &h->c;
    if (fno_steps) *fno_steps = c->fno_steps;
    if (adi_steps) *adi_steps = c->fno_adi_steps;
    if (mean_res_fno) *mean_res_fno = c->fno_steps ? c->fno_res_sum / (double)c->fno_steps : 0.0;
    if (mean_res_adi) *mean_res_adi = c->fno_nres_adi ? c->fno_res_adi_sum / (double)c->fno_nres_adi : 0.0;
    return G4MI_OK;
}

int g4mi_sync(g4mi_ctx *h) {
    if (!h) return G4MI_E_ARG;
    G4_HIP(hipStreamSynchronize(h->c.stream));
    return G4MI_OK;
}

int g4mi_set_timing(g4mi_ctx *h, int on) {
    if (!h) return G4MI_E_ARG;
    Ctx *c = &h->c;
    G4_TRY(fold_timers(c));
    if (on && c->pool.empty()) {
        c->pool.resize((size_t)Ctx::kPoolSteps * Ctx::kEv);
        for (auto &e : c->pool) G4_HIP(hipEventCreate(&e));
    }
    c->timing = on != 0;
    for (double &t : c->t_phase) t = 0;
    c->n_timed = 0;
    return G4MI_OK;
}

int g4mi_kernel_times(g4mi_ctx *h, double *ms, int nslots, long long *nsteps) {
    if (!h || nslots < 0 || nslots > G4MI_NTIMERS || (nslots > 0 && !ms)) return G4MI_E_ARG;
    Ctx *c = &h->c;
    G4_TRY(fold_timers(c));
    const double n = c->n_timed > 0 ? (double)c->n_timed : 1.0;
    for (int k = 0; k < nslots; k++) ms[k] = c->t_phase[k] / n;
    if (nsteps) *nsteps = c->n_timed;
    return G4MI_OK;
}

}  // extern "C"
