// CPU emulation of libchswe's fused kernels -- test infrastructure only (never
// linked into the library).  The kernel sources (swe_fused.inc and the
// headers it needs) are compiled for the host with g++:
//
//   * each block's threads run as fibers (ucontext) that switch at
//     __syncthreads() and at lane shuffles; a barrier releases when every
//     thread of the block has reached it, a shuffle when every lane of the
//     wave has; a thread that is missing from either is reported as a
//     deadlock or a divergent shuffle;
//   * blocks run one at a time, in a seeded random order (no block may
//     depend on another within a launch);
//   * every device buffer is a heap allocation of exactly its size, so that
//     AddressSanitizer (built with -fsanitize=address) reports any access
//     outside it; the padding of the state layout, which the fused kernels
//     never read, holds NaN, so a read there shows up in the results;
//   * results are compared with the CPU oracle (oracle/swe_oracle.c): FP64 at
//     1e-12 relative (max error / max magnitude), FP32 at a stated bar.
//
// The host side mirrors the library's plan set-up and launch sequences
// (coefficient tables, edge records, the stage launches of apply and RK4, the
// multi-instance launches of ensembles, the tile-wise integrals).
//
// usage: swe_emu <case> N order prec [steps] [seed] [flags]
//   case  : apply | rk4 | integrals | diag | tends | latlon | latlonuv | multi
//           | lines (the 1-D point stencils; the only case for order 63)
//   prec  : 64 | 32
//   flags : bit 0 every tile on the general path, bit 1 plain block order,
//           bit 2 compensated summation (rk4)
// Prints one line per field and exits 0 when every field is within the bar.
#include <hip/hip_runtime.h>
#include <setjmp.h>
#include <ucontext.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <random>
#include <string>
#include <vector>

#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/common_interface_defs.h>
#define EMU_ASAN 1
#endif

#include "../../oracle/swe_oracle.h"

namespace emu {

struct D3 {
    unsigned x, y, z;
};
D3 g_block{0, 0, 0}, g_grid{1, 1, 1}, g_bdim{1, 1, 1};

constexpr size_t STACK = 512 * 1024;
enum { RUN = 0, BAR = 1, WAVE = 2, DONE = 3 };
// A fiber starts through makecontext / swapcontext; later switches use
// _setjmp / _longjmp, which skip the signal-mask system call of swapcontext.
struct Fiber {
    ucontext_t ctx;
    jmp_buf jb;
    bool started = false;
    unsigned tid = 0;
    int wait = RUN;
    char *stack = nullptr;
#ifdef EMU_ASAN
    void *fake = nullptr;
#endif
};
std::vector<Fiber> fib;
ucontext_t sched;
jmp_buf sched_jb;
int cur = -1;
const std::function<void()> *body = nullptr;
uint64_t slot[1024];
#ifdef EMU_ASAN
const void *main_bottom = nullptr;
size_t main_size = 0;
void *sched_fake = nullptr;
#endif

[[noreturn]] void die(const char *what)
{
    std::fprintf(stderr, "emu: %s (block %u)\n", what, g_block.x);
    std::exit(3);
}

// flat launches run each thread to completion in turn (kernels without
// barriers or shuffles: the geometry set-up)
bool g_flat = false;
unsigned g_flat_tid = 0;

inline D3 tid() { return D3{g_flat ? g_flat_tid : fib[cur].tid, 0, 0}; }

// fiber -> scheduler
void yield()
{
    Fiber &f = fib[cur];
#ifdef EMU_ASAN
    __sanitizer_start_switch_fiber(&f.fake, main_bottom, main_size);
#endif
    if (!_setjmp(f.jb)) _longjmp(sched_jb, 1);
#ifdef EMU_ASAN
    __sanitizer_finish_switch_fiber(f.fake, &main_bottom, &main_size);
#endif
}

void entry()
{
#ifdef EMU_ASAN
    __sanitizer_finish_switch_fiber(nullptr, &main_bottom, &main_size);
#endif
    (*body)();
    fib[cur].wait = DONE;
#ifdef EMU_ASAN
    __sanitizer_start_switch_fiber(nullptr, main_bottom, main_size);
#endif
    _longjmp(sched_jb, 1);  // never resumed
}

void barrier()
{
    if (g_flat) die("barrier in a flat launch");
    fib[cur].wait = BAR;
    yield();
}

void wave_sync()
{
    if (g_flat) die("lane shuffle in a flat launch");
    fib[cur].wait = WAVE;
    yield();
}

template <class T>
T shfl(T v, int src)
{
    static_assert(sizeof(T) <= 8, "shuffle of at most 8 bytes");
    const unsigned t = fib[cur].tid;
    uint64_t b = 0;
    std::memcpy(&b, &v, sizeof(T));
    slot[t] = b;
    wave_sync();
    T r;
    std::memcpy(&r, &slot[(t & ~63u) | (unsigned)(src & 63)], sizeof(T));
    wave_sync();
    return r;
}

template <class T>
T shfl_down(T v, unsigned d)
{
    const unsigned lane = fib[cur].tid & 63;
    return shfl(v, (int)(lane + d < 64 ? lane + d : lane));
}

void resume(int i)
{
    cur = i;
#ifdef EMU_ASAN
    __sanitizer_start_switch_fiber(&sched_fake, fib[i].stack, STACK);
#endif
    if (!_setjmp(sched_jb)) {
        if (!fib[i].started) {
            fib[i].started = true;
            swapcontext(&sched, &fib[i].ctx);
        } else {
            _longjmp(fib[i].jb, 1);
        }
    }
#ifdef EMU_ASAN
    __sanitizer_finish_switch_fiber(sched_fake, nullptr, nullptr);
#endif
    cur = -1;
}

// run fn as one block of nt threads
void run_block(unsigned nt, const std::function<void()> &fn)
{
    if (nt > 1024 || nt % 64) die("block size must be a multiple of 64, at most 1024");
    if (fib.size() < nt) {
        const size_t old = fib.size();
        fib.resize(nt);
        for (size_t i = old; i < nt; i++) fib[i].stack = (char *)std::malloc(STACK);
    }
    body = &fn;
    for (unsigned i = 0; i < nt; i++) {
        Fiber &f = fib[i];
        getcontext(&f.ctx);
        f.ctx.uc_stack.ss_sp = f.stack;
        f.ctx.uc_stack.ss_size = STACK;
        f.ctx.uc_link = nullptr;
        makecontext(&f.ctx, entry, 0);
        f.tid = i;
        f.wait = RUN;
        f.started = false;
    }
    for (;;) {
        for (unsigned i = 0; i < nt; i++)
            if (fib[i].wait == RUN) resume((int)i);
        // every thread now waits at a barrier or a shuffle, or has finished
        bool released = false;
        for (unsigned w = 0; w < nt / 64; w++) {
            int nwave = 0, ndone = 0;
            for (unsigned l = 0; l < 64; l++) {
                nwave += fib[w * 64 + l].wait == WAVE;
                ndone += fib[w * 64 + l].wait == DONE;
            }
            if (nwave == 0) continue;
            if (nwave + ndone != 64) die("divergent lane shuffle (some lanes of the wave are elsewhere)");
            for (unsigned l = 0; l < 64; l++)
                if (fib[w * 64 + l].wait == WAVE) fib[w * 64 + l].wait = RUN;
            released = true;
        }
        if (released) continue;
        unsigned nbar = 0, ndone = 0;
        for (unsigned i = 0; i < nt; i++) {
            nbar += fib[i].wait == BAR;
            ndone += fib[i].wait == DONE;
        }
        if (ndone == nt) return;
        if (nbar == nt) {
            for (unsigned i = 0; i < nt; i++) fib[i].wait = RUN;
            continue;
        }
        die(ndone ? "barrier not reached by threads that have finished" : "deadlock");
    }
}

std::mt19937_64 g_rng(1);
bool g_shuffle_blocks = true;

// a launch: grid blocks of nt threads in a random order
void launch(unsigned grid, unsigned nt, const std::function<void()> &fn)
{
    std::vector<unsigned> order(grid);
    for (unsigned b = 0; b < grid; b++) order[b] = b;
    if (g_shuffle_blocks) std::shuffle(order.begin(), order.end(), g_rng);
    g_grid = D3{grid, 1, 1};
    g_bdim = D3{nt, 1, 1};
    for (unsigned b : order) {
        g_block = D3{b, 0, 0};
        run_block(nt, fn);
    }
}

// a launch of a kernel without barriers or shuffles
void launch_flat(unsigned grid, unsigned nt, const std::function<void()> &fn)
{
    g_grid = D3{grid, 1, 1};
    g_bdim = D3{nt, 1, 1};
    g_flat = true;
    for (unsigned b = 0; b < grid; b++) {
        g_block = D3{b, 0, 0};
        for (unsigned t = 0; t < nt; t++) {
            g_flat_tid = t;
            fn();
        }
    }
    g_flat = false;
}

// the 24-bit multiplier of the device: both operands must fit in 24 bits
unsigned umul24(unsigned a, unsigned b)
{
    if (a >= (1u << 24) || b >= (1u << 24)) die("__umul24 operand beyond 24 bits");
    return a * b;
}

void release()
{
    for (auto &f : fib) std::free(f.stack);
    fib.clear();
}

}  // namespace emu

// ---- the kernel sources, compiled for the host ---------------------------------
#define threadIdx (emu::tid())
#define blockIdx (emu::g_block)
#define blockDim (emu::g_bdim)
#define gridDim (emu::g_grid)
#define __syncthreads() emu::barrier()
#define __shfl(v, s) emu::shfl((v), (s))
#define __shfl_down(v, d) emu::shfl_down((v), (d))
#define __builtin_amdgcn_s_setprio(x) ((void)0)
#define __umul24(a, b) emu::umul24((a), (b))
#undef __shared__
#define __shared__ static
#undef __launch_bounds__
#define __launch_bounds__(...)
#define amdgpu_waves_per_eu(...)
using std::max;
using std::min;

#include "../../ch_swe_amd/csrc/swe_sbp.hpp"
#include "../../ch_swe_amd/csrc/swe_topology.hpp"
#include "../../ch_swe_amd/csrc/swe_latlon.hpp"
namespace {
#include "../../ch_swe_amd/csrc/swe_geom.hpp"
#include "../../ch_swe_amd/csrc/swe_fused.inc"
#include "../../ch_swe_amd/csrc/swe_fused_host.hpp"
#include "../../ch_swe_amd/csrc/swe_latlon.inc"
}  // namespace

// ---- host side ---------------------------------------------------------------------
namespace {

constexpr double RADIUS = 6.371e6, GRAV = 9.81, HMEAN = 1000.0, OMEGA = 7.292e-5;

template <typename T>
T *dev(std::vector<std::vector<T>> &pool, size_t n, T fill = T(0))
{
    pool.emplace_back(n, fill);
    return pool.back().data();
}

template <typename T>
struct Plan {
    static constexpr int TY = sizeof(T) == 4 ? 16 : 8;
    int N, order;
    Dims d;
    Coef<T> cf{};
    T *rec[2];
    FusedTables ft;
    int flags;
    std::vector<std::vector<T>> pool;

    GeomParams g{};
    Plan(int N_, int order_, int flags_) : N(N_), order(order_), d(make_dims(N_)), flags(flags_)
    {
        const int n = N + 1;
        g.hx = 0.5 * PI / N;
        g.radius = RADIUS;
        g.omega = OMEGA;
        g.axis[0] = 0;
        g.axis[1] = 0;
        g.axis[2] = 1;
        g.coriolis_const = 0;
        g.order = order;
        T *q[4], *tab[8];
        for (auto &p : q) p = dev(pool, n);
        for (auto &p : tab) p = dev(pool, (size_t)n * n);
        T *fj = dev(pool, 6 * (size_t)n * n);
        const unsigned tpb = 256;
        emu::launch_flat((n + tpb - 1) / tpb, tpb, [&] { k_sync_weights<T>(N, g.hx, q[0], q[1], q[2], q[3]); });
        emu::launch_flat((unsigned)((6L * n * n + tpb - 1) / tpb), tpb, [&] {
            k_coef<T>(N, g, tab[0], tab[1], tab[2], tab[3], tab[4], tab[5], tab[6], tab[7], fj);
        });
        cf = Coef<T>{tab[0], tab[1], tab[2], tab[3], tab[4], tab[5], tab[6], tab[7], fj, q[0], q[1], q[2], q[3]};
        for (auto &r : rec) r = dev(pool, (size_t)24 * REC_KINDS * n, T(NAN));
        std::string err;
        build_fused_tables(N, TY, ft, err);
        for (int i = 0; i < ft.nseg; i++) {
            const int4 a = segment_of<TY>(N, i), b = ft.segs[i];
            if (a.x != b.x || a.y != b.y || a.z != b.z || a.w != b.w) emu::die("segment decode");
        }
    }

    // a state: padded layout, NaN outside the interior points; with
    // compensated summation (flag bit 2) followed by its compensation
    // fields, zero
    T *state()
    {
        T *s = dev(pool, (flags & 4 ? 6 : 3) * (size_t)d.F, T(NAN));
        if (flags & 4) std::fill(s + 3 * d.F, s + 6 * d.F, T(0));
        return s;
    }
    void upload(T *st, const double *u, const double *v, const double *h)
    {
        auto put = [&](T *f, const double *src, int nx, int ny) {
            for (int p = 0; p < 6; p++)
                for (int j = 1; j <= ny; j++)
                    for (int i = 1; i <= nx; i++) f[off(d, p, i, j)] = T(src[((size_t)p * ny + (j - 1)) * nx + (i - 1)]);
        };
        put(st, u, N, N + 1);
        put(st + d.F, v, N + 1, N);
        put(st + 2 * d.F, h, N + 1, N + 1);
    }
    void download(const T *st, double *u, double *v, double *h)
    {
        auto get = [&](const T *f, double *dst, int nx, int ny) {
            for (int p = 0; p < 6; p++)
                for (int j = 1; j <= ny; j++)
                    for (int i = 1; i <= nx; i++) dst[((size_t)p * ny + (j - 1)) * nx + (i - 1)] = double(f[off(d, p, i, j)]);
        };
        get(st, u, N, N + 1);
        get(st + d.F, v, N + 1, N);
        get(st + 2 * d.F, h, N + 1, N + 1);
    }

    template <int O>
    StageArgs<T> args(const T *y) const
    {
        StageArgs<T> a{};
        a.d = d;
        a.c = cf;
        a.y = y;
        a.g_hxa = T(GRAV / (0.5 * PI / N * RADIUS));
        a.H0 = T(HMEAN);
        a.ntx = ft.ntx;
        a.nty = ft.nty;
        a.fwx = ft.fwx;
        a.fwy = ft.fwy;
        a.order = O;
        a.xcd = (flags & 2) ? 0 : 1;
        a.general = (flags & 1) ? 1 : 0;
        return a;
    }

    template <int O>
    void records(const T *y, T *r)
    {
        emu::launch(ft.nseg, F_REC_THREADS, [&] { k_records<T, O, TY>(d, cf, y, r); });
    }

    template <int O, int MODE>
    void stage(const StageArgs<T> &a)
    {
        emu::launch(ft.ntiles, F_THREADS, [&] { k_stage<T, O, TY, MODE>(a); });
    }

    // Engine::fused_apply
    template <int O>
    void apply(T *out, const T *in)
    {
        records<O>(in, rec[0]);
        StageArgs<T> a = args<O>(in);
        a.out = out;
        a.rec_in = rec[0];
        stage<O, 0>(a);
    }

    // Engine::fused_rk4
    template <int O>
    void rk4(T *v0, double dt, int nsteps)
    {
        T *acc = state(), *ya = state(), *yb = state();
        T *cmp = (flags & 4) ? v0 + 3 * d.F : nullptr;
        records<O>(v0, rec[0]);
        const T *ins[4] = {v0, ya, yb, ya};
        T *nexts[4] = {ya, yb, ya, nullptr};
        const double A[4] = {0.5 * dt, 0.5 * dt, dt, 0.0};
        const double B[4] = {dt / 6.0, dt / 3.0, dt / 3.0, dt / 6.0};
        for (int it = 0; it < nsteps; it++)
            for (int st = 0; st < 4; st++) {
                StageArgs<T> a = args<O>(ins[st]);
                a.v0 = v0;
                a.acc = acc;
                a.ynext = nexts[st];
                a.out = v0;
                a.cmp = cmp;
                a.a_dt = T(A[st]);
                a.b_dt = T(B[st]);
                a.rec_in = rec[st & 1];
                if (st == 0) stage<O, 1>(a);
                else if (st == 3 && cmp) stage<O, 4>(a);
                else if (st == 3) stage<O, 3>(a);
                else stage<O, 2>(a);
                if (st < 3) records<O>(nexts[st], rec[(st + 1) & 1]);
                else if (it + 1 < nsteps) records<O>(v0, rec[0]);
            }
    }

    // Engine::fused_integrals, the final sum on the host in the same order
    template <int O>
    void integrals(const T *st, double &mass, double &energy)
    {
        const int grid = std::min(ft.ntiles, INT_MAX_BLOCKS);
        std::vector<double> prt(3 * (size_t)grid, NAN);
        double *pp = prt.data();
        emu::launch(grid, F_THREADS, [&] { k_integrals<T, O, TY>(d, cf, st, ft.ntx, ft.nty, ft.fwx, ft.fwy, pp); });
        double s[3] = {0, 0, 0};
        for (int k = 0; k < grid; k++)
            for (int q = 0; q < 3; q++) s[q] += prt[3 * k + q];
        mass = s[0];
        energy = 0.5 * GRAV * s[1] + 0.5 * HMEAN * s[2];
    }

    // Engine::integrals_diag
    template <int O>
    void diag(const T *st, double grav, double out[7])
    {
        const int grid = std::min(ft.ntiles, INT_MAX_BLOCKS);
        std::vector<double> prt(DIAG_K * (size_t)grid, NAN), o(9, NAN);
        double *pp = prt.data(), *op = o.data();
        const double dx = RADIUS * g.hx;
        emu::launch(grid, F_THREADS,
                    [&] { k_diag<T, O, TY>(d, cf, g, st, ft.ntx, ft.nty, ft.fwx, ft.fwy, grav, 1.0 / (dx * dx), pp); });
        emu::launch(1, 64, [&] { k_diag_final(pp, grid, op); });
        for (int k = 0; k < 7; k++) out[k] = o[k];
    }

    // Engine::latlon_h: the plan-time tables, the extended arrays, the
    // interpolation (every buffer sized exactly)
    void latlon(const T *st, int nlat, int nlon, float *out)
    {
        swe::LatlonTables t;
        std::string err;
        if (swe::build_latlon_tables(N, nlat, nlon, t, err)) {
            std::fprintf(stderr, "%s\n", err.c_str());
            std::exit(2);
        }
        const int n = N + 1;
        std::vector<int> col(t.src.size());
        for (size_t k = 0; k < col.size(); k++) {
            const int s = t.src[k], p = s / (n * n), r = s - p * n * n, j = r / n + 1, i = r - (j - 1) * n + 1;
            col[k] = (int)off(d, p, i, j);
        }
        const int nr = (int)t.rp.size() - 1, np = (int)t.base.size();
        std::vector<double> ext(nr, NAN);
        std::vector<float> o(np, NAN);
        const T *h = st + 2 * d.F;
        emu::launch_flat((nr + 255) / 256, 256, [&] {
            k_ll_ext<T>(h, t.rp.data(), col.data(), t.w.data(), nr, ext.data());
        });
        emu::launch_flat((np + 255) / 256, 256, [&] {
            k_ll_interp(ext.data(), t.base.data(), t.fx.data(), t.fy.data(), np, t.ew, o.data());
        });
        std::copy(o.begin(), o.end(), out);
    }

    // Engine::latlon_uv: colocation, the two components' extended arrays,
    // the interpolation and the zonal / meridional turn
    void latlon_uv(const T *st, int nlat, int nlon, float *uo, float *vo)
    {
        swe::LatlonTables t;
        std::string err;
        if (swe::build_latlon_vec_tables(N, nlat, nlon, t, err)) {
            std::fprintf(stderr, "%s\n", err.c_str());
            std::exit(2);
        }
        const int n = N + 1;
        const long nxy = 6L * n * n;
        const int nr = (int)t.rp.size() - 1, np = (int)t.base.size();
        std::vector<double> uv(2 * nxy, NAN), ext(nr, NAN);
        std::vector<float> o(2 * (size_t)np, NAN);
        emu::launch_flat((unsigned)((nxy + 255) / 256), 256, [&] { k_ll_colocate<T>(d, st, st + d.F, uv.data()); });
        emu::launch_flat((nr + 255) / 256, 256, [&] {
            k_ll_ext<double>(uv.data(), t.rp.data(), t.src.data(), t.w.data(), nr, ext.data());
        });
        emu::launch_flat((np + 255) / 256, 256, [&] {
            k_ll_interp_uv(ext.data(), nr / 2, t.base.data(), t.fx.data(), t.fy.data(), t.q.data(), np, t.ew, o.data(),
                           o.data() + np);
        });
        std::copy(o.begin(), o.begin() + np, uo);
        std::copy(o.begin() + np, o.end(), vo);
    }

    // Engine::integral_tends: apply into a tendency state, then the tile sums
    template <int O>
    void tends(T *st, T *tend, double grav, double out[2])
    {
        apply<O>(tend, st);
        const int grid = std::min(ft.ntiles, INT_MAX_BLOCKS);
        std::vector<double> prt(TENDS_K * (size_t)grid, NAN), o(2, NAN);
        double *pp = prt.data(), *op = o.data();
        emu::launch(grid, F_THREADS,
                    [&] { k_tends<T, O, TY>(d, cf, g, st, tend, ft.ntx, ft.nty, ft.fwx, ft.fwy, grav, pp); });
        emu::launch(1, 64, [&] { k_tends_final(pp, grid, op); });
        out[0] = o[0];
        out[1] = o[1];
    }
};

double rel(const std::vector<double> &a, const std::vector<double> &b)
{
    double e = 0, m = 0;
    for (size_t k = 0; k < a.size(); k++) {
        if (!std::isfinite(a[k])) return INFINITY;
        e = std::max(e, std::fabs(a[k] - b[k]));
        m = std::max(m, std::fabs(b[k]));
    }
    return e / std::max(m, 1e-300);
}

struct Fields {
    std::vector<double> u, v, h;
    Fields(int N) : u((size_t)6 * N * (N + 1)), v((size_t)6 * N * (N + 1)), h((size_t)6 * (N + 1) * (N + 1)) {}
};

Fields random_state(int N, unsigned seed)
{
    Fields f(N);
    std::mt19937_64 r(seed);
    std::normal_distribution<double> g(0.0, 1.0);
    for (auto *a : {&f.u, &f.v, &f.h})
        for (double &x : *a) x = g(r);
    return f;
}

swo *oracle(int N, int order)
{
    const double axis[3] = {0, 0, 1};
    return swo_create(N, order, RADIUS, GRAV, HMEAN, OMEGA, axis, 0);
}

int report(const char *what, const Fields &got, const Fields &ref, double bar)
{
    const double e[3] = {rel(got.u, ref.u), rel(got.v, ref.v), rel(got.h, ref.h)};
    int bad = 0;
    const char *nm[3] = {"u", "v", "h"};
    for (int k = 0; k < 3; k++) {
        std::printf("%s %s rel_err %.3e (bar %.1e)\n", what, nm[k], e[k], bar);
        bad |= !(e[k] <= bar);
    }
    return bad;
}

template <typename T, int O>
int run_case(const std::string &cs, int N, int steps, unsigned seed, int flags)
{
    const double bar = sizeof(T) == 8 ? 1e-12 : 2e-5;
    Plan<T> pl(N, O, flags);
    swo *o = oracle(N, O);
    Fields s0 = random_state(N, seed);
    const double dt = 0.5 * (0.5 * PI / N) * RADIUS / std::sqrt(GRAV * HMEAN);  // CFL 0.5
    int bad = 0;
    if (cs == "apply") {
        T *in = pl.state(), *out = pl.state();
        pl.upload(in, s0.u.data(), s0.v.data(), s0.h.data());
        pl.template apply<O>(out, in);
        Fields got(N), ref(N);
        pl.download(out, got.u.data(), got.v.data(), got.h.data());
        swo_apply(o, s0.u.data(), s0.v.data(), s0.h.data(), ref.u.data(), ref.v.data(), ref.h.data());
        bad = report("apply", got, ref, bar);
    } else if (cs == "rk4") {
        T *v0 = pl.state();
        pl.upload(v0, s0.u.data(), s0.v.data(), s0.h.data());
        pl.template rk4<O>(v0, dt, steps);
        Fields got(N), ref = s0;
        pl.download(v0, got.u.data(), got.v.data(), got.h.data());
        swo_rk4(o, ref.u.data(), ref.v.data(), ref.h.data(), dt, steps);
        bad = report("rk4", got, ref, bar);
    } else if (cs == "integrals") {
        T *st = pl.state();
        pl.upload(st, s0.u.data(), s0.v.data(), s0.h.data());
        double m = 0, e = 0;
        pl.template integrals<O>(st, m, e);
        const double mr = swo_mass(o, s0.h.data()), er = swo_energy(o, s0.u.data(), s0.v.data(), s0.h.data());
        double hmax = 0;
        for (double x : s0.h) hmax = std::max(hmax, std::fabs(x));
        const double ibar = sizeof(T) == 8 ? 1e-12 : 1e-6;
        const double em = std::fabs(m - mr) / (hmax * 4 * PI), ee = std::fabs(e / er - 1);
        std::printf("integrals mass err %.3e energy rel err %.3e (bar %.1e)\n", em, ee, ibar);
        bad = !(em <= ibar) || !(ee <= ibar);
    } else if (cs == "diag") {
        T *st = pl.state();
        // a state near rest on a deep layer (the enstrophy divides by h)
        Fields s1 = s0;
        for (double &x : s1.h) x = 500.0 + 10.0 * x;
        pl.upload(st, s1.u.data(), s1.v.data(), s1.h.data());
        double got[7], ref[7];
        pl.template diag<O>(st, 9.80616, got);
        swo_integrals_diag(o, s1.u.data(), s1.v.data(), s1.h.data(), 9.80616, ref);
        const double dbar = sizeof(T) == 8 ? 1e-12 : 2e-6;
        const char *nm[7] = {"hmin", "hmax", "mass", "TE", "KE", "PE", "enstrophy"};
        for (int k = 0; k < 7; k++) {
            const double e = std::fabs(got[k] - ref[k]) / std::max(std::fabs(ref[k]), 1e-300);
            std::printf("diag %s %.17g oracle %.17g rel err %.3e (bar %.1e)\n", nm[k], got[k], ref[k], e, dbar);
            bad |= !(e <= dbar);
        }
    } else if (cs == "latlon") {
        T *st = pl.state();
        pl.upload(st, s0.u.data(), s0.v.data(), s0.h.data());
        const int nlat = 2 * N + 1, nlon = 4 * N;
        std::vector<float> got((size_t)nlat * nlon), ref((size_t)nlat * nlon);
        pl.latlon(st, nlat, nlon, got.data());
        std::vector<double> hh = s0.h;
        if (sizeof(T) == 4)
            for (double &x : hh) x = (double)(float)x;  // the oracle reads what the plan holds
        swo_latlon_h(o, hh.data(), nlat, nlon, ref.data());
        // float outputs of FP64 sums that differ in their last bits
        int ulps = 0;
        for (size_t k = 0; k < got.size(); k++) {
            int32_t a, b;
            std::memcpy(&a, &got[k], 4);
            std::memcpy(&b, &ref[k], 4);
            ulps = std::max(ulps, std::abs(a - b));
            if (!std::isfinite(got[k])) ulps = 1 << 30;
        }
        std::printf("latlon %d x %d max ulp difference %d (bar 1)\n", nlat, nlon, ulps);
        bad = ulps > 1;
    } else if (cs == "latlonuv") {
        T *st = pl.state();
        pl.upload(st, s0.u.data(), s0.v.data(), s0.h.data());
        const int nlat = 2 * N + 1, nlon = 4 * N;
        const size_t np = (size_t)nlat * nlon;
        std::vector<float> gu(np), gv(np), ru(np), rv(np);
        pl.latlon_uv(st, nlat, nlon, gu.data(), gv.data());
        std::vector<double> uu = s0.u, vv = s0.v;
        if (sizeof(T) == 4) {
            for (double &x : uu) x = (double)(float)x;
            for (double &x : vv) x = (double)(float)x;
        }
        swo_latlon_uv(o, uu.data(), vv.data(), nlat, nlon, ru.data(), rv.data());
        // the frame change of the halo is derived differently here and in
        // the oracle: FP64 sums that differ in their last bits, rounded to
        // float and turned into zonal / meridional components
        double mx = 0, e = 0;
        for (size_t k = 0; k < np; k++) {
            mx = std::max(mx, (double)std::max(std::fabs(ru[k]), std::fabs(rv[k])));
            e = std::max(e, (double)std::max(std::fabs(gu[k] - ru[k]), std::fabs(gv[k] - rv[k])));
            if (!std::isfinite(gu[k]) || !std::isfinite(gv[k])) e = INFINITY;
        }
        std::printf("latlonuv %d x %d max error / max value %.3e (bar 4e-7)\n", nlat, nlon, e / mx);
        bad = !(e <= 4e-7 * mx);
    } else if (cs == "tends") {
        T *st = pl.state(), *tend = pl.state();
        Fields s1 = s0;
        for (double &x : s1.h) x = 500.0 + 10.0 * x;
        pl.upload(st, s1.u.data(), s1.v.data(), s1.h.data());
        double got[2], ref[4];
        pl.template tends<O>(st, tend, 9.80616, got);
        swo_integral_tends(o, s1.u.data(), s1.v.data(), s1.h.data(), 9.80616, ref);
        // measured against the sum of the terms' absolute values: both cancel
        const double tbar = sizeof(T) == 8 ? 1e-12 : 2e-6;
        const char *nm[2] = {"TE_tend", "Enstrophy_tend"};
        for (int k = 0; k < 2; k++) {
            const double e = std::fabs(got[k] - ref[k]) / ref[2 + k];
            std::printf("tends %s %.17g oracle %.17g err/scale %.3e (bar %.1e)\n", nm[k], got[k], ref[k], e, tbar);
            bad |= !(e <= tbar);
        }
    } else {
        std::fprintf(stderr, "unknown case %s\n", cs.c_str());
        bad = 2;
    }
    swo_destroy(o);
    return bad;
}

// Ensemble::run for two members of one order, N and N2, each with its own
// step count and dt (multi-instance kernels over (member, tile) work lists)
template <typename T, int O>
int run_multi(int N, int N2, int steps, unsigned seed)
{
    constexpr int TY = Plan<T>::TY;
    const double bar = sizeof(T) == 8 ? 1e-12 : 2e-5;
    const int Ns[2] = {N, N2}, ns[2] = {steps, std::max(1, steps - 1)};
    Plan<T> p0(N, O, 0), p1(N2, O, 0);
    Plan<T> *pl[2] = {&p0, &p1};
    T *v0[2], *acc[2], *ya[2], *yb[2], *r0[2], *r1[2];
    Fields s0[2] = {random_state(N, seed), random_state(N2, seed + 1)};
    double dt[2];
    std::vector<std::vector<T>> pool;
    for (int i = 0; i < 2; i++) {
        v0[i] = pl[i]->state();
        pl[i]->upload(v0[i], s0[i].u.data(), s0[i].v.data(), s0[i].h.data());
        acc[i] = pl[i]->state();
        ya[i] = pl[i]->state();
        yb[i] = pl[i]->state();
        r0[i] = dev(pool, (size_t)24 * REC_KINDS * (Ns[i] + 1), T(NAN));
        r1[i] = dev(pool, (size_t)24 * REC_KINDS * (Ns[i] + 1), T(NAN));
        dt[i] = (0.4 + 0.1 * i) * (0.5 * PI / Ns[i]) * RADIUS / std::sqrt(GRAV * HMEAN);
    }
    // args[stage][member] as Ensemble::member_args
    std::vector<StageArgs<T>> args(8);
    for (int s = 0; s < 4; s++)
        for (int i = 0; i < 2; i++) {
            const T *ins[4] = {v0[i], ya[i], yb[i], ya[i]};
            T *nexts[4] = {ya[i], yb[i], ya[i], nullptr};
            const double A[4] = {0.5 * dt[i], 0.5 * dt[i], dt[i], 0.0};
            const double B[4] = {dt[i] / 6.0, dt[i] / 3.0, dt[i] / 3.0, dt[i] / 6.0};
            StageArgs<T> a = pl[i]->template args<O>(ins[s]);
            a.v0 = v0[i];
            a.acc = acc[i];
            a.ynext = nexts[s];
            a.out = v0[i];
            a.a_dt = T(A[s]);
            a.b_dt = T(B[s]);
            a.rec_in = (s & 1) ? r1[i] : r0[i];
            a.rec_out = (s & 1) ? r0[i] : r1[i];
            args[(size_t)s * 2 + i] = a;
        }
    // members ordered by step count (member 0 runs longer); work lists per member
    std::vector<int2> wl, rl;
    int wpre[3] = {0, 0, 0}, rpre[3] = {0, 0, 0};
    for (int i = 0; i < 2; i++) {
        for (int t = 0; t < pl[i]->ft.ntiles; t++) wl.push_back(make_int2(i, tile_at(t, pl[i]->ft.ntx, pl[i]->ft.nty)));
        for (int sg = 0; sg < pl[i]->ft.nseg; sg++) rl.push_back(make_int2(i, sg));
        wpre[i + 1] = (int)wl.size();
        rpre[i + 1] = (int)rl.size();
    }
    const StageArgs<T> *A0 = args.data();
    const int2 *W = wl.data(), *R = rl.data();
    emu::launch(rpre[2], F_REC_THREADS, [&] { k_records_multi<T, O, TY, 0>(A0, R); });
    for (int it = 0; it < ns[0]; it++) {
        const int k = it < ns[1] ? 2 : 1;  // members still running
        const int nw = wpre[k], nr = rpre[k];
        for (int s = 0; s < 4; s++) {
            const StageArgs<T> *a = args.data() + (size_t)s * 2;
            if (s == 0) emu::launch(nw, F_THREADS, [&] { k_stage_multi<T, O, TY, 1>(a, W, nw); });
            else if (s == 3) emu::launch(nw, F_THREADS, [&] { k_stage_multi<T, O, TY, 3>(a, W, nw); });
            else emu::launch(nw, F_THREADS, [&] { k_stage_multi<T, O, TY, 2>(a, W, nw); });
            if (s < 3) emu::launch(nr, F_REC_THREADS, [&] { k_records_multi<T, O, TY, 1>(a, R); });
            else emu::launch(nr, F_REC_THREADS, [&] { k_records_multi<T, O, TY, 2>(a, R); });
        }
    }
    int bad = 0;
    for (int i = 0; i < 2; i++) {
        swo *o = oracle(Ns[i], O);
        Fields got(Ns[i]), ref = s0[i];
        pl[i]->download(v0[i], got.u.data(), got.v.data(), got.h.data());
        swo_rk4(o, ref.u.data(), ref.v.data(), ref.h.data(), dt[i], ns[i]);
        char nm[64];
        std::snprintf(nm, sizeof nm, "multi member %d (N=%d, %d steps)", i, Ns[i], ns[i]);
        bad |= report(nm, got, ref, bar);
        swo_destroy(o);
    }
    return bad;
}

// The 1-D SBP stencils of swe_sbp.hpp (the multi-kernel path's point
// functions, SAT terms included) on random lines of n = N .. N + 3 centres with
// random halo values, against the oracle's line operators.
template <typename T, int O>
int run_lines(int N, unsigned seed)
{
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    const double bar = sizeof(T) == 8 ? 1e-13 : 2e-6;
    int bad = 0;
    for (int n = N; n < N + 4; n++) {
        const int len = n + 1 + 2 * 4;  // index 1 - 4 .. n + 1 + 4
        std::vector<double> in(len), ref(n + 1);
        std::vector<T> tin(len);
        for (int k = 0; k < len; k++) {
            in[k] = U(rng);
            tin[k] = T(in[k]);
            in[k] = double(tin[k]);
        }
        const T *f0 = tin.data() + 4 - 1;  // index 0
        static const char *names[4] = {"d_i2c", "d_c2i", "i_i2c", "i_c2i"};
        for (int op = 0; op < 4; op++) {
            swo_line_op(O, op, n, in.data(), ref.data(), 1);
            const int nout = (op == 0 || op == 2) ? n : n + 1;
            double err = 0, mag = 0;
            for (int k = 1; k <= nout; k++) {
                auto f = swe::line(f0, 1);
                T g = op == 0   ? swe::d_i2c_pt<O, T>(f, 0, n, k)
                      : op == 1 ? swe::d_c2i_pt<O, T>(f, 0, n, k)
                      : op == 2 ? swe::i_i2c_pt<O, T>(f, 0, n, k)
                                : swe::i_c2i_pt<O, T>(f, 0, n, k);
                err = std::max(err, std::fabs(double(g) - ref[k - 1]));
                mag = std::max(mag, std::fabs(ref[k - 1]));
            }
            const double rel = err / std::max(mag, 1e-300);
            std::printf("  %-6s sbp%d n=%d rel %.3e%s\n", names[op], O, n, rel, rel > bar ? "  <-- FAIL" : "");
            bad |= rel > bar;
        }
    }
    return bad;
}

template <typename T>
int dispatch(const std::string &cs, int N, int order, int steps, unsigned seed, int flags)
{
    if (cs == "lines")
        return order == 21 ? run_lines<T, 21>(N, seed) : order == 42 ? run_lines<T, 42>(N, seed)
                                                                    : run_lines<T, 63>(N, seed);
    if (order == 63) {
        std::fprintf(stderr, "sbp63 runs on the multi-kernel path: only the lines case applies\n");
        return 1;
    }
    if (cs == "multi") {
        // the second member: a different N with the same order
        const int N2 = N + 7;
        return order == 21 ? run_multi<T, 21>(N, N2, steps, seed) : run_multi<T, 42>(N, N2, steps, seed);
    }
    return order == 21 ? run_case<T, 21>(cs, N, steps, seed, flags) : run_case<T, 42>(cs, N, steps, seed, flags);
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s apply|rk4|integrals|diag|tends|latlon|latlonuv|multi|lines N order 64|32 [steps] [seed] [flags]\n", argv[0]);
        return 2;
    }
    const std::string cs = argv[1];
    const int N = std::atoi(argv[2]), order = std::atoi(argv[3]), prec = std::atoi(argv[4]);
    const int steps = argc > 5 ? std::atoi(argv[5]) : 1;
    const unsigned seed = argc > 6 ? (unsigned)std::atoi(argv[6]) : 1u;
    const int flags = argc > 7 ? std::atoi(argv[7]) : 0;
    if ((order != 21 && order != 42 && order != 63) || (prec != 64 && prec != 32) ||
        N < (order == 21 ? 4 : order == 42 ? 8 : 12)) {
        std::fprintf(stderr, "bad arguments\n");
        return 2;
    }
    emu::g_rng.seed(seed * 7919u + 17u);
    std::printf("case %s N=%d order=%d fp%d steps=%d seed=%u flags=%d\n", cs.c_str(), N, order, prec, steps, seed,
                flags);
    const int bad = prec == 64 ? dispatch<double>(cs, N, order, steps, seed, flags)
                               : dispatch<float>(cs, N, order, steps, seed, flags);
    emu::release();
    std::printf("%s\n", bad ? "FAIL" : "ok");
    return bad ? 1 : 0;
}
