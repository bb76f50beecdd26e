// kernels.hip -- hand-written gfx950 kernels of the phase2 state-vector hot
// path and the thin C ABI that launches them (declared in include/phase2_hip.h).
//
// Register layout: 2^nqb complex128 amplitudes, interleaved (re, im), 16 B
// each, in one hipMalloc buffer; basis index i has qubit k at bit k.
//
// Arithmetic contract (bit identity with the reference CPU backend,
// phase2/qreg_qreg.c:149-160): built with -ffp-contract=off and every
// product/sum goes through mul()/add() below, so each component update is
// round(round(c*x) + round(+-s*y)) with c, s computed by the host libm.
// Negations and component swaps are exact.  Nothing here is FMA-contracted.
//
// Kernels:
//   k_rot_pair    plain pass, one thread per amplitude pair (i, i^x), i<j
//   k_rot_diag    plain pass for diagonal terms (x == 0)
//   k_rot_fused   one HBM pass applying a run of terms in an LDS tile that is
//                 a coset of span{low bits, flip masks of the run}
//   k_fill        zeroing (no memset nodes, so the call is graph-safe)
//   k_gather      amp[idx[k]] for the multidet overlap
//   k_sc_scatter  Slater-Condon coefficients into the register
//   k_sc_inner    Slater-Condon inner product, wavefront-shuffle reduction
//   k_norm2       sum |a|^2, same reduction
//   k_reduce2     second stage of the two-stage reductions (fixed order)

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdarg>
#include <algorithm>

#include "phase2_hip.h"

// Staging of ph2hip_gather_noalloc: PH2HIP_GATHER_NOALLOC_MAX indices, then
// as many amplitudes (16 B each).  Each device holds its own copy.  In host
// code this name is a host-side stand-in, not a device address: the address
// comes from hipGetSymbolAddress at ph2hip_init (g_scratch below).
static __device__ __align__(256) uint64_t scratch[3 * PH2HIP_GATHER_NOALLOC_MAX];

namespace {

// ---------------------------------------------------------------------------
// error state and stream
// ---------------------------------------------------------------------------

char g_err[256] = "";
hipStream_t g_stream = nullptr;
int g_device = -1;
uint64_t g_persist_wgs = 512;   // fused-pass workgroups (2 per CU by default)
uint32_t g_ncu = 256;          // compute units of the device
double *g_red = nullptr;       // reduction partials (device)
size_t g_red_cap = 0;
// device address of `scratch` per device ordinal, taken once at ph2hip_init
constexpr int kMaxDevices = 64;
void *g_scratch[kMaxDevices] = {};

int fail(const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(g_err, sizeof g_err, fmt, ap);
	va_end(ap);
	return -1;
}

#define HIPCHK(call)                                                        \
	do {                                                                \
		hipError_t e_ = (call);                                     \
		if (e_ != hipSuccess)                                       \
			return fail("%s: %s", #call, hipGetErrorString(e_)); \
	} while (0)

int ensure_init()
{
	if (g_stream)
		return 0;
	return ph2hip_init(-1);
}

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ double mul(double a, double b)
{
	return __dmul_rn(a, b);
}

__device__ __forceinline__ double add(double a, double b)
{
	return __dadd_rn(a, b);
}

__device__ __forceinline__ uint64_t insert_zero(uint64_t t, uint32_t h)
{
	const uint64_t lo = t & ((UINT64_C(1) << h) - 1);
	return ((t >> h) << (h + 1)) | lo;
}

__device__ __forceinline__ uint32_t insert_zero32(uint32_t t, uint32_t h)
{
	const uint32_t lo = t & ((1u << h) - 1);
	return ((t >> h) << (h + 1)) | lo;
}

// Update of one pair (a_i, a_j), i < j, P|i> = i^r4 |j>:
//   a_i <- c a_i + i conj(s i^r4) a_j,   a_j <- c a_j + i (s i^r4) a_i.
// For odd r4 the partner enters as itself, for even r4 as i*partner
// = (-im, re); the sign of s is - for r4 in {2,3} (a_i) and {1,2} (a_j).
__device__ __forceinline__ void rot_pair(double2 &x, double2 &y, int r4,
	double c, double s)
{
	const bool odd = r4 & 1;
	const double si = (r4 & 2) ? -s : s;
	const double sj = ((r4 + 1) & 2) ? -s : s;
	const double pr = odd ? y.x : -y.y, pi = odd ? y.y : y.x;
	const double qr = odd ? x.x : -x.y, qi = odd ? x.y : x.x;
	const double2 nx = make_double2(add(mul(c, x.x), mul(si, pr)),
		add(mul(c, x.y), mul(si, pi)));
	const double2 ny = make_double2(add(mul(c, y.x), mul(sj, qr)),
		add(mul(c, y.y), mul(sj, qi)));
	x = nx;
	y = ny;
}

// Diagonal term (x == 0): a <- c a + i (s z) a with z = (-1)^parity.
__device__ __forceinline__ void rot_diag(double2 &a, int odd_parity,
	double c, double s)
{
	const double sj = odd_parity ? -s : s;
	a = make_double2(add(mul(c, a.x), mul(sj, -a.y)),
		add(mul(c, a.y), mul(sj, a.x)));
}

constexpr int kBlock = 256;
// HIP limits gridDim.x * blockDim.x to below 2^32; rotation kernels loop
// over the work with at most this many workgroups.
constexpr uint32_t kMaxGrid = 1u << 22;

// ---------------------------------------------------------------------------
// plain pass
// ---------------------------------------------------------------------------

// Pair t -> i = t with a zero inserted at bit h (the top bit of x), j = i^x
// > i.  Consecutive lanes take consecutive t, so both streams coalesce.
template <bool NT>
__device__ __forceinline__ void st16(double2 *p, double2 v)
{
	if (NT) {
		__builtin_nontemporal_store(v.x, &p->x);
		__builtin_nontemporal_store(v.y, &p->y);
	} else {
		*p = v;
	}
}

template <int UNR, bool NT>
__global__ __launch_bounds__(kBlock) void k_rot_pair(
	double2 *__restrict__ a, uint64_t npairs, uint64_t x, uint64_t z,
	uint32_t h, uint32_t y, double c, double s)
{
	for (uint64_t chunk = blockIdx.x; chunk * (kBlock * UNR) < npairs;
		chunk += gridDim.x) {
	const uint64_t t0 = chunk * (kBlock * UNR) + threadIdx.x;
	double2 vi[UNR], vj[UNR];
	uint64_t ii[UNR];
#pragma unroll
	for (int u = 0; u < UNR; u++) {
		const uint64_t t = t0 + (uint64_t)u * kBlock;
		ii[u] = insert_zero(t, h);
		if (t < npairs) {
			vi[u] = a[ii[u]];
			vj[u] = a[ii[u] ^ x];
		}
	}
#pragma unroll
	for (int u = 0; u < UNR; u++) {
		const uint64_t t = t0 + (uint64_t)u * kBlock;
		if (t < npairs) {
			const int par = __popcll(ii[u] & z) & 1;
			rot_pair(vi[u], vj[u], (int)((y + 2 * par) & 3), c, s);
			st16<NT>(&a[ii[u]], vi[u]);
			st16<NT>(&a[ii[u] ^ x], vj[u]);
		}
	}
	}
}

template <int UNR, bool NT>
__global__ __launch_bounds__(kBlock) void k_rot_diag(
	double2 *__restrict__ a, uint64_t namp, uint64_t z, double c, double s)
{
	for (uint64_t chunk = blockIdx.x; chunk * (kBlock * UNR) < namp;
		chunk += gridDim.x) {
	const uint64_t i0 = chunk * (kBlock * UNR) + threadIdx.x;
	double2 v[UNR];
#pragma unroll
	for (int u = 0; u < UNR; u++) {
		const uint64_t i = i0 + (uint64_t)u * kBlock;
		if (i < namp)
			v[u] = a[i];
	}
#pragma unroll
	for (int u = 0; u < UNR; u++) {
		const uint64_t i = i0 + (uint64_t)u * kBlock;
		if (i < namp) {
			rot_diag(v[u], __popcll(i & z) & 1, c, s);
			st16<NT>(&a[i], v[u]);
		}
	}
	}
}

// ---------------------------------------------------------------------------
// fused pass
// ---------------------------------------------------------------------------
//
// Tile: the coset  base ^ span{e_0..e_{L-1}, w_0..w_{g-1}}  of 2^TB
// amplitudes (TB = L + g), held in LDS.  The w_b have zero low L bits,
// distinct leading bits p_b (ascending in b) and zeros at every other p_b'.
// Tile coordinate v: bits [0,L) are index bits 0..L-1, bit L+b selects w_b;
// the tile index spreads over the index bits that are neither low nor
// pivots.  In these coordinates a term with flip mask x has xt, and
//   parity(i & z) = parity(base & z) ^ parity(v & zt),   i < i^x <=> v < v^xt,
// so the plain update runs unchanged on v.
//
// Inside the tile the terms of the pass are cut into micro-groups whose xt
// span at most D dimensions.  A group has its own basis u_0..u_{D-1} (tile
// coordinates, same echelon form); each thread owns one coset
//   bu ^ span{u}  (2^D amplitudes, in registers),
// applies every term of the group to it without communication, and writes
// it back: one LDS round trip per group instead of one per term.  Within a
// coset the term's flip pattern xtp (D bits) is a compile-time case, so
// register indices stay static.
//
// LDS slots are XOR-swizzled (bits >= 4 folded into bits 0..3) so that the
// strided cosets of a wave spread over the banks.

// The descriptors are kernel arguments read with wave-uniform indices.
// Every field the kernels read is an aligned 32-bit word (byte fields are
// read through the word that holds them), so all of them come in through
// scalar loads: a byte or unaligned read would become a vector-memory load
// that shares the wait counter with the tile's HBM traffic.
struct alignas(16) FusedTerm {
	uint64_t z;    // physical phase mask (per-tile parity)
	double c, s;
	uint32_t zt;   // phase mask in tile coordinates (bits 0..15) and the
	               // per-register signs parity(w & ztp) (bits 16..31)
	union {
		struct {
			uint8_t xtp;   // flip mask in group coordinates
			uint8_t ztp;   // parity(u_i & zt), group coordinates
			uint8_t y;     // popcount(x & z) & 3
			uint8_t pad;
		};
		uint32_t meta;
	};
};

struct alignas(16) FusedGroup {
	union {
		uint16_t u[4];   // group basis, tile coordinates
		uint32_t uw[2];
	};
	union {
		uint8_t piv[4];  // its leading bits, ascending
		uint32_t pivw;
	};
	union {
		uint8_t jpos[12];// tile bit of each coset-index bit (the free bits)
		uint32_t jposw[3];
	};
	union {
		struct {
			// terms first .. first + count; for a half-tile pass,
			// the exchange to the next group splits the tile by tile
			// bit xh, carried by coset-index bit xk
			uint8_t first, count, xh, xk;
		};
		uint32_t fcw;
	};
	uint32_t pad3[1];
};

constexpr int kMaxPer = (1 << PH2HIP_TILE_MAX) / kBlock;

struct FusedPass {
	uint64_t w[PH2HIP_TILE_MAX];   // gathered basis vectors
	uint64_t roff[kMaxPer];        // gathered offset of load/store step m
	// Offset (row part ^ column part) of coset element w of the first
	// ([0]) and last ([1]) micro-group, for groups that move their cosets
	// straight between HBM and registers (dflags).
	uint64_t dcw[2][16];
	uint32_t piv[PH2HIP_TILE_MAX]; // their leading bits, ascending
	union {
		struct {
			uint8_t L, g, ngroups, nterms;
		};
		uint32_t hdr;
	};
	// bit 0: the first group loads its cosets from HBM (the tile is not
	// staged in LDS first); bit 1: the last group stores them to HBM.
	uint32_t dflags;
	FusedGroup grp[PH2HIP_PASS_MAX];
	FusedTerm t[PH2HIP_PASS_MAX];
};

// Device-side reads of the packed fields.
__device__ __forceinline__ uint32_t pass_L(const FusedPass &P)
{
	return P.hdr & 0xffu;
}
__device__ __forceinline__ uint32_t pass_g(const FusedPass &P)
{
	return (P.hdr >> 8) & 0xffu;
}
__device__ __forceinline__ uint32_t pass_ngroups(const FusedPass &P)
{
	return (P.hdr >> 16) & 0xffu;
}

static_assert(sizeof(FusedPass) <= 4000, "kernel argument too large");

// LDS slot of tile coordinate v: bits >= 4 folded into bits 0..3.  The map
// is GF(2)-linear, so slot(a ^ b) = slot(a) ^ slot(b).
typedef double v2d __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t lds_slot(uint32_t v)
{
	return v ^ (((v >> 4) ^ (v >> 8)) & 15u);
}

__device__ __forceinline__ double flip_sign(double v, uint32_t f)
{
	return __hiloint2double(__double2hiint(v) ^ (int)(f << 31),
		__double2loint(v));
}

// One pair update in place: (x, y) of one coset pair, c uniform, si the
// per-thread signed sine, si = sb * sg with sg = +-1.0 uniform.  ODD:
// x <- c x + si y,  y <- c y - si x;  even:  x <- c x + si (i y),
// y <- c y + si (i x), per component as in rot_pair.  Each result is
// round(round(c u) +- round(si v)), the same roundings as
// add(mul(c, u), mul(+-si, v)) (negation is exact).  Written as one block
// with the four values read and written in place: as plain code the
// compiler gives each term's results new registers and then copies the
// whole coset back once per term at the loop back-edge.  The real and the
// imaginary halves are two independent 2x2 updates sharing four
// temporaries, and si is formed inside the block, so a pair holds five
// temporaries at most (the scheduler would otherwise form the signed
// sines of all pairs of a term early and keep them live).
template <bool ODD>
__device__ __forceinline__ void pair_inplace(double &xr, double &xi,
	double &yr, double &yi, double c, double sb, double sg)
{
	double si, t0, t1, t2, t3;
	if (ODD)
		__asm__(
			"v_mul_f64 %[si], %[sg], %[sb]\n\t"
			"v_mul_f64 %[t0], %[c], %[xr]\n\t"
			"v_mul_f64 %[t2], %[c], %[yr]\n\t"
			"v_mul_f64 %[t1], %[si], %[yr]\n\t"
			"v_mul_f64 %[t3], %[si], %[xr]\n\t"
			"v_add_f64 %[xr], %[t0], %[t1]\n\t"
			"v_add_f64 %[yr], %[t2], -%[t3]\n\t"
			"v_mul_f64 %[t0], %[c], %[xi]\n\t"
			"v_mul_f64 %[t2], %[c], %[yi]\n\t"
			"v_mul_f64 %[t1], %[si], %[yi]\n\t"
			"v_mul_f64 %[t3], %[si], %[xi]\n\t"
			"v_add_f64 %[xi], %[t0], %[t1]\n\t"
			"v_add_f64 %[yi], %[t2], -%[t3]"
			: [xr] "+v"(xr), [xi] "+v"(xi), [yr] "+v"(yr),
			  [yi] "+v"(yi), [si] "=&v"(si), [t0] "=&v"(t0),
			  [t1] "=&v"(t1), [t2] "=&v"(t2), [t3] "=&v"(t3)
			: [c] "s"(c), [sb] "v"(sb), [sg] "s"(sg));
	else
		__asm__(
			"v_mul_f64 %[si], %[sg], %[sb]\n\t"
			"v_mul_f64 %[t0], %[c], %[xr]\n\t"
			"v_mul_f64 %[t2], %[c], %[yi]\n\t"
			"v_mul_f64 %[t1], %[si], %[yi]\n\t"
			"v_mul_f64 %[t3], %[si], %[xr]\n\t"
			"v_add_f64 %[xr], %[t0], -%[t1]\n\t"
			"v_add_f64 %[yi], %[t2], %[t3]\n\t"
			"v_mul_f64 %[t0], %[c], %[xi]\n\t"
			"v_mul_f64 %[t2], %[c], %[yr]\n\t"
			"v_mul_f64 %[t1], %[si], %[yr]\n\t"
			"v_mul_f64 %[t3], %[si], %[xi]\n\t"
			"v_add_f64 %[xi], %[t0], %[t1]\n\t"
			"v_add_f64 %[yr], %[t2], -%[t3]"
			: [xr] "+v"(xr), [xi] "+v"(xi), [yr] "+v"(yr),
			  [yi] "+v"(yi), [si] "=&v"(si), [t0] "=&v"(t0),
			  [t1] "=&v"(t1), [t2] "=&v"(t2), [t3] "=&v"(t3)
			: [c] "s"(c), [sb] "v"(sb), [sg] "s"(sg));
}

// Diagonal update in place: a <- c a + sj (i a) with sj = sb * sg, i.e.
// (c re - sj im, c im + sj re), the roundings of rot_diag.
__device__ __forceinline__ void diag_inplace(double &re, double &im,
	double c, double sb, double sg)
{
	double sj, t0, t1, t2, t3;
	__asm__(
		"v_mul_f64 %[sj], %[sg], %[sb]\n\t"
		"v_mul_f64 %[t0], %[c], %[re]\n\t"
		"v_mul_f64 %[t2], %[c], %[im]\n\t"
		"v_mul_f64 %[t1], %[sj], %[im]\n\t"
		"v_mul_f64 %[t3], %[sj], %[re]\n\t"
		"v_add_f64 %[re], %[t0], -%[t1]\n\t"
		"v_add_f64 %[im], %[t2], %[t3]"
		: [re] "+v"(re), [im] "+v"(im), [sj] "=&v"(sj), [t0] "=&v"(t0),
		  [t1] "=&v"(t1), [t2] "=&v"(t2), [t3] "=&v"(t3)
		: [c] "s"(c), [sb] "v"(sb), [sg] "s"(sg));
}

// High word of +-1.0 for coset register w: -1.0 if bit w of the term's
// sign mask is set.  A bit test and a select on the scalar unit.
__device__ __forceinline__ double reg_sign(uint32_t sgn, int w)
{
	return __hiloint2double((sgn >> w & 1u) ? (int)0xbff00000 :
		0x3ff00000, 0);
}

// +-1.0 with the sign bit f (an exact multiplier)
__device__ __forceinline__ double unit_sign(uint32_t f)
{
	return __hiloint2double(0x3ff00000 | (int)(f << 31), 0);
}

// One term with compile-time flip pattern XT on a coset of 2^D amplitudes
// held as (re, im) register quads, so that the 16-byte HBM and LDS moves
// read and write them in place (with re and im in separate arrays the
// compiler pairs them up with two copies per move).  r4 = y + 2*par (mod 4)
// with par = bpt ^ parity(w & ztp).  Odd y: the partner enters as itself
// and s_j = -s_i; even y: as i*partner and s_j = s_i.  s_i = s with its
// sign flipped by par ^ bit1(y); sb already carries bpt ^ bit1(y), so per
// pair only parity(w & ztp) (uniform) flips.
template <int D, int XT, bool ODD>
__device__ __forceinline__ void coset_term(v2d (&a)[1 << D], uint32_t sgn,
	double sb, double c)
{
	constexpr int H = 31 - __builtin_clz(XT);
#pragma unroll
	for (int q = 0; q < (1 << (D - 1)); q++) {
		const int w = ((q >> H) << (H + 1)) | (q & ((1 << H) - 1));
		double xr = a[w].x, xi = a[w].y, yr = a[w ^ XT].x,
			yi = a[w ^ XT].y;
		// per pair only the uniform sign (scalar unit) changes: one
		// exact multiply by +-1.0 held in scalar registers
		pair_inplace<ODD>(xr, xi, yr, yi, c, sb, reg_sign(sgn, w));
		a[w].x = xr;
		a[w].y = xi;
		a[w ^ XT].x = yr;
		a[w ^ XT].y = yi;
	}
}

// xtp is in 1 .. 2^D - 1 (built so on the host); the last pattern takes
// every value not matched before it, so that no path leaves the coset
// unchanged: such a path makes the compiler copy all 2^(D+1) coset
// registers once per term at the loop back-edge.
template <int D, bool ODD, int XT = 1>
__device__ __forceinline__ void coset_dispatch(v2d (&a)[1 << D],
	uint32_t xtp, uint32_t sgn, double sb, double c)
{
	if constexpr (XT == (1 << D) - 1) {
		coset_term<D, XT, ODD>(a, sgn, sb, c);
	} else {
		if (xtp == XT) {
			coset_term<D, XT, ODD>(a, sgn, sb, c);
			return;
		}
		coset_dispatch<D, ODD, XT + 1>(a, xtp, sgn, sb, c);
	}
}

__device__ __forceinline__ uint64_t tile_base(const FusedPass &P, uint64_t t)
{
	uint64_t base = t << pass_L(P);
	const uint32_t g = pass_g(P);
	for (uint32_t b = 0; b < g; b++)
		base = insert_zero(base, P.piv[b]);
	return base;
}

// The terms of one micro-group (P.t[first .. first + count), fcw) on one
// coset of 2^D registers with base coordinate bu in the tile at base.
template <int D>
__device__ __forceinline__ void group_terms(v2d (&a)[1 << D],
	const FusedPass &P, uint32_t fcw, uint32_t bu, uint64_t base)
{
	const uint32_t kfirst = fcw & 0xffu;
	const uint32_t kend = kfirst + ((fcw >> 8) & 0xffu);
	// term k + 1 is loaded (scalar loads) while term k runs
	uint64_t tz = P.t[kfirst].z;
	double c = P.t[kfirst].c, ts = P.t[kfirst].s;
	uint32_t tzt = P.t[kfirst].zt, meta = P.t[kfirst].meta;
	for (uint32_t k = kfirst; k < kend; k++) {
		const uint32_t kn = k + 1 < kend ? k + 1 : k;
		const uint64_t tz_n = P.t[kn].z;
		const double c_n = P.t[kn].c, ts_n = P.t[kn].s;
		const uint32_t tzt_n = P.t[kn].zt,
			meta_n = P.t[kn].meta;
		const uint32_t xtp = meta & 0xffu;
		const uint32_t ty = (meta >> 16) & 0xffu;
		const uint32_t sgn = tzt >> 16;
		const uint32_t bpt = (__popc(bu & (tzt & 0xffffu)) ^
			__popcll(base & tz)) & 1u;
		if (xtp == 0) {
			// diagonal: a <- c a + s_j (i a), s_j = +-s
			const double sb = flip_sign(ts, bpt);
#pragma unroll
			for (int w = 0; w < (1 << D); w++) {
				double re = a[w].x, im = a[w].y;
				diag_inplace(re, im, c, sb, reg_sign(sgn, w));
				a[w].x = re;
				a[w].y = im;
			}
		} else {
			const double sb = flip_sign(ts,
				bpt ^ ((ty >> 1) & 1u));
			if (ty & 1)
				coset_dispatch<D, true>(a, xtp, sgn, sb,
					c);
			else
				coset_dispatch<D, false>(a, xtp, sgn, sb,
					c);
		}
		tz = tz_n;
		c = c_n;
		ts = ts_n;
		tzt = tzt_n;
		meta = meta_n;
		// each quad stays one register quad across terms: without this
		// the D = 3 coset was copied whole at the loop back-edge
#pragma unroll
		for (int w = 0; w < (1 << D); w++)
			__asm__ volatile("" : "+v"(a[w]));
	}
}

// Tile coordinate of coset j's base: j's bits spread over the free tile
// bits jpos (three packed words).
template <int TB, int D>
__device__ __forceinline__ uint32_t coset_base(uint32_t j, uint32_t jw0,
	uint32_t jw1, uint32_t jw2)
{
	const uint32_t jw[3] = { jw0, jw1, jw2 };
	uint32_t bu = 0;
#pragma unroll
	for (int k = 0; k < TB - D; k++)
		bu |= (j >> k & 1u) << ((jw[k >> 2] >> (8 * (k & 3))) & 0xffu);
	return bu;
}

// HBM index of tile coordinate bu (row part ^ column part).
__device__ __forceinline__ uint64_t hbm_index(const FusedPass &P,
	uint64_t base, uint32_t bu, uint32_t L, uint32_t g)
{
	uint64_t abu = base ^ (bu & ((1u << L) - 1));
	for (uint32_t b = 0; b < g; b++)
		if ((bu >> (L + b)) & 1u)
			abu ^= P.w[b];
	return abu;
}

// All micro-groups of a pass on one tile.  Each group's cosets come from
// and go back to the LDS tile, except that with dflags bit 0 the first
// group reads its cosets straight from HBM (the tile was not staged) and
// with bit 1 the last group writes them straight to HBM (the tile is not
// stored from LDS afterwards).  Those groups map coset-index bits 0..3 to
// tile bits 0..3 (checked on the host, with L >= 4), so 16 lanes move one
// 256-byte run; a pass of one group then never touches LDS.
template <int TB, int D>
__device__ __forceinline__ void tile_groups(v2d *__restrict__ a, v2d *tile,
	const FusedPass &P, uint64_t base, uint32_t tid, uint32_t dflags)
{
	constexpr uint32_t NSUB = (1u << TB) >> D;    // cosets per group
	static_assert(NSUB % kBlock == 0, "cosets must cover the block");
	const uint32_t ngroups = pass_ngroups(P);
	const uint32_t L = pass_L(P), g = pass_g(P);
	for (uint32_t gi = 0; gi < ngroups; gi++) {
		const uint32_t uw0 = P.grp[gi].uw[0], uw1 = P.grp[gi].uw[1];
		const uint32_t jw0 = P.grp[gi].jposw[0], jw1 = P.grp[gi].jposw[1],
			jw2 = P.grp[gi].jposw[2];
		const uint32_t fcw = P.grp[gi].fcw;
		const uint32_t gu[4] = { uw0 & 0xffffu, uw0 >> 16, uw1 & 0xffffu,
			uw1 >> 16 };
		// byte offsets of the basis vectors' LDS slots (the slot map is
		// linear, so an element's offset is the XOR of these)
		uint32_t gs[4];
#pragma unroll
		for (int i = 0; i < 4; i++)
			gs[i] = lds_slot(gu[i]) << 4;
		const bool gload = gi == 0 && (dflags & 1u);
		const bool gstore = gi + 1 == ngroups && (dflags & 2u);
		if (!gload)
			__syncthreads();
		for (uint32_t j = tid; j < NSUB; j += kBlock) {
			const uint32_t bu = coset_base<TB, D>(j, jw0, jw1, jw2);
			// byte offset of the coset base's slot: the per-element
			// offsets (uniform) are XORed in pre-shifted, one VALU
			// operation per LDS access
			uint32_t sbu16 = lds_slot(bu) << 4;
			__asm__("" : "+v"(sbu16));   // keeps the shift out of
			                             // the per-access XOR
			const uint64_t abu = (gload || gstore) ?
				hbm_index(P, base, bu, L, g) : 0;
			v2d cs[1 << D];
			if (gload) {
#pragma unroll
				for (int w = 0; w < (1 << D); w++)
					cs[w] = a[abu ^ P.dcw[0][w]];
			} else {
#pragma unroll
				for (int w = 0; w < (1 << D); w++) {
					uint32_t cw16 = 0;
#pragma unroll
					for (int i = 0; i < D; i++)
						if (w >> i & 1)
							cw16 ^= gs[i];
					cs[w] = *reinterpret_cast<const v2d *>(
						reinterpret_cast<const char *>(tile) +
						(sbu16 ^ cw16));
				}
			}
			group_terms<D>(cs, P, fcw, bu, base);
			if (gstore) {
#pragma unroll
				for (int w = 0; w < (1 << D); w++)
					a[abu ^ P.dcw[1][w]] = cs[w];
			} else {
#pragma unroll
				for (int w = 0; w < (1 << D); w++) {
					uint32_t cw16 = 0;
#pragma unroll
					for (int i = 0; i < D; i++)
						if (w >> i & 1)
							cw16 ^= gs[i];
					*reinterpret_cast<v2d *>(
						reinterpret_cast<char *>(tile) +
						(sbu16 ^ cw16)) = cs[w];
				}
			}
		}
	}
	if (!(dflags & 2u))
		__syncthreads();
}

// Slot of tile coordinate v in a buffer of half a tile: bits h and TB-1
// exchanged, then bits 4 and up folded into bits 0..3 through
// y ^ (y >> 1) ^ (y >> 3) with y = v >> 4, then bit TB-1 dropped.  Linear,
// and one-to-one on each half {v : bit h of v = const}.
//
// Banks, as measured on gfx950 with the LDS conflict counters: a 16-byte
// write serves 8 lanes per clock (lane bits 0..2) over 32 banks, so those
// lanes need distinct slots mod 8; a 16-byte read serves 16 lanes per
// clock (lane bits 0..3) over 64 banks, distinct slots mod 16.  The fold
// gives tile bits 4..10 the bank vectors 1, 3, 6, 0xd, 0xa, 4, 8 (bits 11
// and 12 none), so that the host (half_bank) finds, for nearly every
// group, three free lane bits independent mod 8 and a fourth independent
// mod 16.  The whole-tile fold (bank vector t mod 4 for tile bit t) runs
// out of a residue for many groups: 15% of LDS cycles were conflicts at
// tile 12 with it, 6% at tile 13.
template <int TB>
__device__ __forceinline__ uint32_t half_slot(uint32_t v, uint32_t h)
{
	const uint32_t t = ((v >> h) ^ (v >> (TB - 1))) & 1u;
	v ^= t * ((1u << h) | (1u << (TB - 1)));
	v ^= ((v >> 4) ^ (v >> 5) ^ (v >> 7)) & 15u;
	return v & ((1u << (TB - 1)) - 1);
}

// Byte offsets in the half buffer of a group's basis vectors (uniform).
template <int TB>
__device__ __forceinline__ void half_offsets(const FusedGroup &G, uint32_t h,
	uint32_t (&gs)[4])
{
	const uint32_t uw0 = G.uw[0], uw1 = G.uw[1];
	gs[0] = half_slot<TB>(uw0 & 0xffffu, h) << 4;
	gs[1] = half_slot<TB>(uw0 >> 16, h) << 4;
	gs[2] = half_slot<TB>(uw1 & 0xffffu, h) << 4;
	gs[3] = half_slot<TB>(uw1 >> 16, h) << 4;
}

// All micro-groups of a pass on one tile with half a tile of LDS, from the
// first group's cosets in cs (coset base bu of group 0, read from HBM by
// the caller) to the last group's, which go back to HBM.  Between groups q
// and q+1 the cosets move through the buffer in two rounds, one per half of
// the tile (split bit xh of group q; the threads of one half are the waves
// whose coset-index bit xk is that half).  One coset per thread: 2^(TB - D)
// threads.  fresh: the buffer has no earlier readers in this workgroup (its
// first tile), so the first exchange needs no barrier before writing.
template <int TB, int D>
__device__ __forceinline__ void half_groups(v2d *__restrict__ a, v2d *buf,
	const FusedPass &P, uint64_t base, uint32_t tid, v2d (&cs)[1 << D],
	uint32_t bu, bool fresh)
{
	// the split index bits (the top two) must select whole 64-lane waves
	static_assert(TB - D >= 8 && TB - D <= 10, "2^(TB-D) threads, 256..1024");
	const uint32_t ngroups = pass_ngroups(P);
	const uint32_t L = pass_L(P), g = pass_g(P);
	for (uint32_t gi = 0;; gi++) {
		const uint32_t fcw = P.grp[gi].fcw;
		group_terms<D>(cs, P, fcw, bu, base);
		if (gi + 1 == ngroups)
			break;
		const FusedGroup &N = P.grp[gi + 1];
		const uint32_t h = (fcw >> 16) & 0xffu, xk = fcw >> 24;
		const uint32_t half = __builtin_amdgcn_readfirstlane(
			(tid >> xk) & 1u);
		// the thread index is re-read per exchange: hoisted out of
		// the loop, its bit extractions would hold 8 VGPRs throughout
		uint32_t jt = tid;
		__asm__ volatile("" : "+v"(jt));
		const uint32_t bn = coset_base<TB, D>(jt, N.jposw[0],
			N.jposw[1], N.jposw[2]);
		uint32_t gw[4], gr[4];
		half_offsets<TB>(P.grp[gi], h, gw);
		half_offsets<TB>(N, h, gr);
		uint32_t sw = half_slot<TB>(bu, h) << 4;
		uint32_t sr = half_slot<TB>(bn, h) << 4;
		__asm__("" : "+v"(sw), "+v"(sr));
		for (uint32_t q = 0; q < 2; q++) {
			// the buffer's previous readers are done (it has none before
			// the first exchange of a workgroup's first tile)
			if (gi > 0 || q > 0 || !fresh)
				__syncthreads();
			if (half == q) {
#pragma unroll
				for (int w = 0; w < (1 << D); w++) {
					uint32_t c16 = 0;
#pragma unroll
					for (int i = 0; i < D; i++)
						if (w >> i & 1)
							c16 ^= gw[i];
					*reinterpret_cast<v2d *>(
						reinterpret_cast<char *>(buf) +
						(sw ^ c16)) = cs[w];
				}
			}
			__syncthreads();
			if (half == q) {
#pragma unroll
				for (int w = 0; w < (1 << D); w++) {
					uint32_t c16 = 0;
#pragma unroll
					for (int i = 0; i < D; i++)
						if (w >> i & 1)
							c16 ^= gr[i];
					cs[w] = *reinterpret_cast<const v2d *>(
						reinterpret_cast<const char *>(buf) +
						(sr ^ c16));
				}
			}
		}
		bu = bn;
	}
	const uint64_t abu = hbm_index(P, base, bu, L, g);
#pragma unroll
	for (int w = 0; w < (1 << D); w++)
		a[abu ^ P.dcw[1][w]] = cs[w];
}

// One tile in the half-tile form: the first group's cosets from HBM, then
// half_groups.
template <int TB, int D>
__device__ __forceinline__ void tile_groups_half(v2d *__restrict__ a,
	v2d *buf, const FusedPass &P, uint64_t base, uint32_t tid)
{
	const uint32_t bu = coset_base<TB, D>(tid, P.grp[0].jposw[0],
		P.grp[0].jposw[1], P.grp[0].jposw[2]);
	const uint64_t abu = hbm_index(P, base, bu, pass_L(P), pass_g(P));
	v2d cs[1 << D];
#pragma unroll
	for (int w = 0; w < (1 << D); w++)
		cs[w] = a[abu ^ P.dcw[0][w]];
	half_groups<TB, D>(a, buf, P, base, tid, cs, bu, true);
}

// Direct HBM->LDS tile loads (gfx950 global_load_lds_dwordx4): the data
// does not pass through registers, so all of a thread's loads of the tile
// are in flight at once whatever the register pressure of the rotation
// code.  Lane l of wave w in load step m fills LDS slot s = m*256 + 64w + l,
// i.e. tile coordinate lds_slot(s) = s ^ f with f = ((s >> 4) ^ (s >> 8)) &
// 15 (the swizzle is an involution).  f only touches coordinate bits 0..3,
// which lie inside a row when L >= 4, so the lane reads column (s ^ f) &
// lmask of the row of s: every 16 lanes still cover one 256-byte run.
// Requires L >= 4.
#define PH2_LDS(p) ((__attribute__((address_space(3))) void *)(p))

// s_waitcnt with vmcnt <= N and no wait on the other counters (gfx9
// encoding: vmcnt bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8).
template <uint32_t N>
__device__ __forceinline__ void vm_wait()
{
	static_assert(N < 64, "vmcnt field");
	__asm__ volatile("" ::: "memory");
	__builtin_amdgcn_s_waitcnt((N & 15u) | ((N >> 4) << 14) | (7u << 4) |
		(15u << 8));
	__asm__ volatile("" ::: "memory");
}

template <int TB>
__device__ __forceinline__ void dma_tile(const v2d *__restrict__ a,
	v2d *__restrict__ buf,
	const FusedPass &P, uint64_t bo, uint32_t tid)
{
	constexpr uint32_t PER = (1u << TB) / kBlock;
	const uint32_t lmask = (1u << pass_L(P)) - 1;
	const uint32_t wave0 = __builtin_amdgcn_readfirstlane(tid & ~63u);
#pragma unroll
	for (uint32_t m = 0; m < PER; m++) {
		const uint32_t u = tid + m * kBlock;
		const uint32_t f = ((tid >> 4) ^ m) & 15u;
		__builtin_amdgcn_global_load_lds(
			(const void *)(a + (bo ^ P.roff[m] ^ ((u ^ f) & lmask))),
			PH2_LDS(buf + m * kBlock + wave0), 16, 0, 0);
	}
}

template <int TB>
__device__ __forceinline__ void store_tile(v2d *__restrict__ a,
	const v2d *buf, const FusedPass &P, uint64_t bo, uint32_t tid)
{
	constexpr uint32_t PER = (1u << TB) / kBlock;
	const uint32_t lmask = (1u << pass_L(P)) - 1;
#pragma unroll
	for (uint32_t m = 0; m < PER; m++)
		a[bo ^ P.roff[m] ^ ((tid + m * kBlock) & lmask)] =
			buf[lds_slot(tid + m * kBlock)];
}

// Load one tile from HBM into LDS, apply the pass, write it back.
template <int TB, int D>
__device__ __forceinline__ void fused_tile(v2d *__restrict__ a,
	v2d *tile, const FusedPass &P, uint64_t t)
{
	constexpr uint32_t PER = (1u << TB) / kBlock;
	const uint32_t L = pass_L(P), lmask = (1u << L) - 1, tid = threadIdx.x;
	const uint32_t g = pass_g(P);
	uint64_t wt = 0;   // gathered offset of this thread's row part
	for (uint32_t b = 0; b < g; b++)
		if ((tid >> L) >> b & 1)
			wt ^= P.w[b];
	const uint64_t base = tile_base(P, t);
	const uint64_t bo = base ^ wt;
	const uint32_t dflags = P.dflags;
	if (dflags & 1u) {
		// the first group reads its cosets from HBM itself
	} else if (L >= 4) {
		dma_tile<TB>(a, tile, P, bo, tid);
		vm_wait<0>();
	} else {
		// staged through registers, a few loads in flight at a time
#pragma unroll 4
		for (uint32_t m = 0; m < PER; m++)
			tile[lds_slot(tid + m * kBlock)] =
				a[bo ^ P.roff[m] ^ ((tid + m * kBlock) & lmask)];
	}
	tile_groups<TB, D>(a, tile, P, base, tid, dflags);
	if (!(dflags & 2u))
		store_tile<TB>(a, tile, P, bo, tid);
}

// Waves per SIMD that the LDS of the whole-tile form allows (160 KiB per
// CU, 2^TB x 16 B per 256-thread workgroup), given to the compiler as its
// register target: left free it spends registers that cost occupancy (114
// VGPRs at tile 11, where five workgroups fit the LDS).
constexpr int lds_waves(int TB)
{
	// 160 KiB / (2^TB * 16 B) workgroups, one wave per SIMD each; tiles
	// 12 and 13 fit at most two workgroups, so no target there
	return TB > 11 ? 1 : (10240 >> TB) < 8 ? (10240 >> TB) : 8;
}

// Workgroups are dealt to the 8 XCDs round-robin (workgroup w to XCD
// w mod 8).  Remap so that each XCD takes a contiguous range of tiles:
// workgroup w = 8 l + x runs tile start(x) + l, where XCD x owns
// q + (x < r) tiles (q = ntiles / 8, r = ntiles mod 8).  A bijection on
// [0, ntiles); measured 4.5% faster per fused pass at 33 qubits than tile w.
__device__ __forceinline__ uint64_t xcd_tile(uint64_t w, uint64_t ntiles)
{
	const uint64_t q = ntiles >> 3, r = ntiles & 7, x = w & 7, l = w >> 3;
	return x * q + (x < r ? x : r) + l;
}

// One tile per workgroup (grid = number of tiles); HBM latency is hidden
// by the co-resident workgroups of the CU.
template <int TB, int D>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(lds_waves(TB))))
void k_rot_fused(
	double2 *__restrict__ amp, const FusedPass P, uint64_t ntiles)
{
	__shared__ v2d tile[1u << TB];
	// 2-D grid when the tiles outnumber kMaxGrid (gridDim.x * 256 < 2^32)
	const uint64_t w = (uint64_t)blockIdx.y * gridDim.x + blockIdx.x;
	if (w >= ntiles)
		return;           // whole workgroup: no barrier is left waiting
	const uint64_t t = xcd_tile(w, ntiles);
	// clang vector type: plain 16-byte accesses that stay in registers
	fused_tile<TB, D>(reinterpret_cast<v2d *>(amp), tile, P, t);
}

// Half-tile form of k_rot_fused (descriptor flag 4): half the LDS, so
// twice the workgroups per CU.  One coset per thread: 256 threads at tile
// 12, 512 at tile 13.
template <int TB, int D>
__global__ __launch_bounds__((1 << TB) >> D) void k_rot_fused_half(
	double2 *__restrict__ amp, const FusedPass P, uint64_t ntiles)
{
	__shared__ v2d buf[1u << (TB - 1)];
	const uint64_t w = (uint64_t)blockIdx.y * gridDim.x + blockIdx.x;
	if (w >= ntiles)
		return;           // whole workgroup: no barrier is left waiting
	const uint64_t t = xcd_tile(w, ntiles);
	tile_groups_half<TB, D>(reinterpret_cast<v2d *>(amp), buf, P,
		tile_base(P, t), threadIdx.x);
}

// Persistent half-tile form with the next tile's first-group cosets loaded
// into registers while the current tile is rotated: PF of the 2^D
// amplitudes of each coset (the rest are loaded when the tile starts), so
// each workgroup keeps HBM reads in flight through its compute phase
// instead of only at its start.  Workgroup b walks virtual workgroups
// b, b + G, b + 2G, ... (G = gridDim.x, a multiple of 8, so each stays on
// its XCD and the XCD's workgroups take neighbouring tiles, as xcd_tile
// orders them).  Every workgroup leaves the loop once its next virtual
// index reaches ntiles.
template <int TB, int D, int PF>
__global__ __launch_bounds__((1 << TB) >> D)
void k_rot_fused_hpipe(double2 *__restrict__ amp, const FusedPass P,
	uint64_t ntiles)
{
	static_assert(PF > 0 && PF <= (1 << D), "prefetched amplitudes");
	__shared__ v2d buf[1u << (TB - 1)];
	v2d *__restrict__ a = reinterpret_cast<v2d *>(amp);
	const uint32_t tid = threadIdx.x;
	uint64_t v = blockIdx.x;
	if (v >= ntiles)
		return;           // whole workgroup: no barrier is left waiting
	const uint32_t bu0 = coset_base<TB, D>(tid, P.grp[0].jposw[0],
		P.grp[0].jposw[1], P.grp[0].jposw[2]);
	// HBM offset of the first group's coset base within any tile
	// (hbm_index is base ^ a function of bu)
	const uint64_t off0 = hbm_index(P, 0, bu0, pass_L(P), pass_g(P));
	uint64_t base = tile_base(P, xcd_tile(v, ntiles));
	v2d cs[1 << D], pre[PF];
#pragma unroll
	for (int w = 0; w < PF; w++)
		pre[w] = a[(base ^ off0) ^ P.dcw[0][w]];
	for (bool fresh = true;; fresh = false) {
#pragma unroll
		for (int w = 0; w < PF; w++)
			cs[w] = pre[w];
#pragma unroll
		for (int w = PF; w < (1 << D); w++)
			cs[w] = a[(base ^ off0) ^ P.dcw[0][w]];
		const uint64_t vn = v + gridDim.x;
		const bool more = vn < ntiles;
		const uint64_t bn = more ? tile_base(P, xcd_tile(vn, ntiles)) : 0;
		if (more) {
#pragma unroll
			for (int w = 0; w < PF; w++)
				pre[w] = a[(bn ^ off0) ^ P.dcw[0][w]];
		}
		half_groups<TB, D>(a, buf, P, base, tid, cs, bu0, fresh);
		if (!more)
			break;
		v = vn;
		base = bn;
	}
}

// Batched pass: pass p of many registers in one launch.  Work item i covers
// tiles [tile0_i, tile0_{i+1}) of register i (amplitudes at amp_off_i) with
// descriptor desc_i; a workgroup finds its item by a uniform binary search.
struct BatchItem {
	uint64_t amp_off;
	uint64_t tile0;
	uint64_t desc;
};

// HALF: every descriptor of the launch has the half-tile form (half the
// LDS and one coset per thread, as k_rot_fused_half).
template <int TB, int D, bool HALF = false>
__global__ __launch_bounds__(HALF ? (1 << TB) >> D : kBlock)
__attribute__((amdgpu_waves_per_eu(HALF ? 1 : lds_waves(TB))))
void k_rot_fused_batch(
	double2 *__restrict__ amp, const FusedPass *__restrict__ descs,
	const BatchItem *__restrict__ items, uint32_t nitems, uint64_t ntiles)
{
	__shared__ v2d tile[1u << (HALF ? TB - 1 : TB)];
	// tile order kept round-robin over the XCDs here: the registers of a
	// batch differ in work per tile, and interleaving balances the XCDs
	// (contiguous ranges measured 8% slower on the 512-register batch)
	const uint64_t t = (uint64_t)blockIdx.y * gridDim.x + blockIdx.x;
	if (t >= ntiles)
		return;
	uint32_t lo = 0, hi = nitems;
	while (hi - lo > 1) {
		const uint32_t mid = (lo + hi) / 2;
		if (items[mid].tile0 <= t)
			lo = mid;
		else
			hi = mid;
	}
	const BatchItem it = items[lo];
	if constexpr (HALF) {
		const FusedPass &P = descs[it.desc];
		tile_groups_half<TB, D>(reinterpret_cast<v2d *>(amp) + it.amp_off,
			tile, P, tile_base(P, t - it.tile0), threadIdx.x);
	} else {
		fused_tile<TB, D>(reinterpret_cast<v2d *>(amp) + it.amp_off, tile,
			descs[it.desc], t - it.tile0);
	}
}

// Persistent, software-pipelined pass: each workgroup walks tiles
// blockIdx.x, +gridDim.x, ...; the next tile's HBM reads are in flight in
// registers while the current tile is rotated in LDS, and its writes are
// issued without waiting.  The groups always go through LDS here.
template <int TB, int D>
__global__ __launch_bounds__(kBlock) void k_rot_fused_pipe(
	double2 *__restrict__ amp, const FusedPass P, uint64_t ntiles)
{
	constexpr uint32_t PER = (1u << TB) / kBlock;
	__shared__ v2d tile[1u << TB];
	v2d *__restrict__ a = reinterpret_cast<v2d *>(amp);
	const uint32_t L = pass_L(P), lmask = (1u << L) - 1, tid = threadIdx.x;
	const uint32_t g = pass_g(P);
	uint64_t wt = 0;
	for (uint32_t b = 0; b < g; b++)
		if ((tid >> L) >> b & 1)
			wt ^= P.w[b];
	v2d pre[PER];
	uint64_t t = blockIdx.x;
	if (t < ntiles) {
		const uint64_t bn = tile_base(P, t) ^ wt;
#pragma unroll
		for (uint32_t m = 0; m < PER; m++)
			pre[m] = a[bn ^ P.roff[m] ^ ((tid + m * kBlock) & lmask)];
	}
	for (; t < ntiles; t += gridDim.x) {
		const uint64_t base = tile_base(P, t);
		__syncthreads();          // previous tile fully written out
#pragma unroll
		for (uint32_t m = 0; m < PER; m++)
			tile[lds_slot(tid + m * kBlock)] = pre[m];
		const uint64_t tn = t + gridDim.x;
		if (tn < ntiles) {
			const uint64_t bn = tile_base(P, tn) ^ wt;
#pragma unroll
			for (uint32_t m = 0; m < PER; m++)
				pre[m] = a[bn ^ P.roff[m] ^
					((tid + m * kBlock) & lmask)];
		}
		tile_groups<TB, D>(a, tile, P, base, tid, 0u);
		const uint64_t bo = base ^ wt;
#pragma unroll
		for (uint32_t m = 0; m < PER; m++)
			a[bo ^ P.roff[m] ^ ((tid + m * kBlock) & lmask)] =
				tile[lds_slot(tid + m * kBlock)];
	}
}

// ---------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_fill(double2 *__restrict__ a,
	uint64_t n)
{
	const uint64_t stride = (uint64_t)gridDim.x * kBlock;
	for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
		i += stride)
		a[i] = make_double2(0.0, 0.0);
}

__global__ __launch_bounds__(kBlock) void k_gather(
	const double2 *__restrict__ a, const uint64_t *__restrict__ idx,
	double2 *__restrict__ out, uint64_t n)
{
	for (uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k < n;
		k += (uint64_t)gridDim.x * kBlock)
		out[k] = a[idx[k]];
}

__global__ __launch_bounds__(kBlock) void k_scatter_set(
	double2 *__restrict__ a, const uint64_t *__restrict__ idx,
	const double2 *__restrict__ val, uint64_t n)
{
	for (uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k < n;
		k += (uint64_t)gridDim.x * kBlock)
		a[idx[k]] = val[k];
}

// splitmix64 output function of x (Steele, Lea and Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014): a bijective 64-bit mix.
__device__ __forceinline__ uint64_t mix64(uint64_t x)
{
	x = (x ^ (x >> 30)) * UINT64_C(0xbf58476d1ce4e5b9);
	x = (x ^ (x >> 27)) * UINT64_C(0x94d049bb133111eb);
	return x ^ (x >> 31);
}

// A double in [-1, 1) from the top 53 bits of h: k * 2^-52 - 1, exact.
__device__ __forceinline__ double unit_of(uint64_t h)
{
	return add(mul((double)(h >> 11), 0x1p-52), -1.0);
}

// amp[i] = (unit_of(mix64(seed + (2i+1) g)), unit_of(mix64(seed + (2i+2) g)))
// with g the golden-ratio increment: a reproducible dense state for checks
// at any width without a host copy (the host restates it in numpy).
__global__ __launch_bounds__(kBlock) void k_fill_hash(
	double2 *__restrict__ a, uint64_t n, uint64_t seed)
{
	const uint64_t g = UINT64_C(0x9e3779b97f4a7c15);
	const uint64_t stride = (uint64_t)gridDim.x * kBlock;
	for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
		i += stride)
		a[i] = make_double2(unit_of(mix64(seed + (2 * i + 1) * g)),
			unit_of(mix64(seed + (2 * i + 2) * g)));
}

// Number of components (re and im counted separately) where a and b differ
// as IEEE doubles (-0 == +0, NaN unequal to everything): per-thread counts,
// a wave sum, one atomic add per wave.
__global__ __launch_bounds__(kBlock) void k_count_neq(
	const double2 *__restrict__ a, const double2 *__restrict__ b, uint64_t n,
	unsigned long long *__restrict__ count)
{
	const uint64_t stride = (uint64_t)gridDim.x * kBlock;
	uint32_t c = 0;
	for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
		i += stride) {
		const double2 u = a[i], v = b[i];
		c += (u.x != v.x) + (u.y != v.y);
	}
	unsigned long long t = c;
#pragma unroll
	for (int off = 32; off > 0; off >>= 1)
		t += __shfl_xor(t, off, 64);
	if ((threadIdx.x & 63) == 0 && t)
		atomicAdd(count, t);
}

// ---------------------------------------------------------------------------
// device-side sampler (qDRIFT and composite draws)
// ---------------------------------------------------------------------------

// One stream of draws: xoshiro256** state, its CDF table (offset into the
// concatenated tables), the number of draws and where they go.
struct DrawJob {
	uint64_t s[4];
	uint64_t cdf_off, len, ndraw, out_off;
};

constexpr uint32_t kDrawChunk = 1024;   // uniforms staged in LDS per round

__device__ __forceinline__ uint64_t rotl64(uint64_t x, int k)
{
	return (x << k) | (x >> (64 - k));
}

// Index of the cumulative table y[0 .. len) that the uniform u selects:
// the binary descent over halving strides, advancing while y[i + d] <= u,
// then the linear walk while y[i] <= u short of the last entry
// (lib/prob.c:67-78; the same comparisons, so ties resolve the same way).
__device__ __forceinline__ uint64_t cdf_pick(const double *__restrict__ y,
	uint64_t len, double u)
{
	uint64_t i = 0;
	for (uint64_t d = len >> 1; d > 0; d >>= 1)
		if (y[i + d] <= u)
			i += d;
	while (i + 1 < len && y[i] <= u)
		i++;
	return i;
}

// One 64-lane workgroup per stream.  The generator is one sequential
// stream, so every lane steps the same (wave-uniform, scalar-unit) state and
// lane 0 stages each uniform (x >> 11) * 2^-53 in LDS
// (lib/xoshiro256ss.c:72-85, include/xoshiro256ss.h:58-59); the lanes then
// invert the table for 64 draws at a time.
__global__ __launch_bounds__(64) void k_cdf_draw(DrawJob *__restrict__ jobs,
	const double *__restrict__ cdf, uint32_t *__restrict__ out)
{
	__shared__ double u[kDrawChunk];
	const uint32_t lane = threadIdx.x;
	DrawJob *J = jobs + blockIdx.x;
	uint64_t s0 = J->s[0], s1 = J->s[1], s2 = J->s[2], s3 = J->s[3];
	const uint64_t len = J->len, ndraw = J->ndraw, o = J->out_off;
	const double *y = cdf + J->cdf_off;
	for (uint64_t b = 0; b < ndraw; b += kDrawChunk) {
		const uint32_t n = (uint32_t)(ndraw - b < kDrawChunk ?
			ndraw - b : kDrawChunk);
		for (uint32_t d = 0; d < n; d++) {
			const uint64_t r = rotl64(s1 * 5, 7) * 9;
			const uint64_t t = s1 << 17;
			s2 ^= s0;
			s3 ^= s1;
			s1 ^= s2;
			s0 ^= s3;
			s2 ^= t;
			s3 = rotl64(s3, 45);
			if (lane == 0)
				u[d] = (double)(r >> 11) * 0x1.0p-53;
		}
		__syncthreads();
		for (uint32_t d = lane; d < n; d += 64)
			out[o + b + d] = (uint32_t)cdf_pick(y, len, u[d]);
		__syncthreads();
	}
	if (lane == 0) {
		J->s[0] = s0;
		J->s[1] = s1;
		J->s[2] = s2;
		J->s[3] = s3;
	}
}

__device__ __forceinline__ uint64_t taper_idx(uint64_t v, uint32_t n_sites)
{
	const uint64_t lo = (v >> 1) & ((UINT64_C(1) << (n_sites - 1)) - 1);
	return lo | ((v >> (n_sites + 1)) << (n_sites - 1));
}

__global__ __launch_bounds__(kBlock) void k_sc_scatter(
	double2 *__restrict__ a, const uint64_t *__restrict__ ma,
	const double *__restrict__ da, const uint64_t *__restrict__ mb,
	const double *__restrict__ db, uint64_t Ma, uint64_t Mb,
	uint32_t n_sites, double weight, int tapered, int accumulate)
{
	const uint64_t total = Ma * Mb;
	const uint64_t stride = (uint64_t)gridDim.x * kBlock;
	for (uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
		k < total; k += stride) {
		const uint64_t ia = k / Mb, ib = k - ia * Mb;
		const double cf = mul(mul(weight, da[ia]), db[ib]);
		if (fabs(cf) < 1.0e-12)
			continue;
		uint64_t idx = ma[ia] | (mb[ib] << n_sites);
		if (tapered)
			idx = taper_idx(idx, n_sites);
		if (accumulate)
			a[idx].x = add(a[idx].x, cf);
		else
			a[idx] = make_double2(cf, 0.0);
	}
}

// Wavefront (64 lanes) sum via cross-lane shuffles, then the 4 waves of the
// block through LDS; thread 0 returns the block total.
__device__ __forceinline__ double2 block_sum2(double2 v)
{
	__shared__ double2 part[kBlock / 64];
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) {
		v.x = add(v.x, __shfl_xor(v.x, off, 64));
		v.y = add(v.y, __shfl_xor(v.y, off, 64));
	}
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	if (lane == 0)
		part[wave] = v;
	__syncthreads();
	double2 r = make_double2(0.0, 0.0);
	if (threadIdx.x == 0)
		for (int w = 0; w < kBlock / 64; w++)
			r = make_double2(add(r.x, part[w].x), add(r.y, part[w].y));
	return r;
}

__global__ __launch_bounds__(kBlock) void k_sc_inner(
	const double2 *__restrict__ a, const uint64_t *__restrict__ ma,
	const double *__restrict__ da, const uint64_t *__restrict__ mb,
	const double *__restrict__ db, uint64_t Ma, uint64_t Mb,
	uint32_t n_sites, double weight, int tapered,
	double2 *__restrict__ partial)
{
	const uint64_t total = Ma * Mb;
	const uint64_t stride = (uint64_t)gridDim.x * kBlock;
	double2 acc = make_double2(0.0, 0.0);
	for (uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
		k < total; k += stride) {
		const uint64_t ia = k / Mb, ib = k - ia * Mb;
		const double cf = mul(mul(weight, da[ia]), db[ib]);
		if (fabs(cf) < 1.0e-12)
			continue;
		uint64_t idx = ma[ia] | (mb[ib] << n_sites);
		if (tapered)
			idx = taper_idx(idx, n_sites);
		const double2 v = a[idx];
		acc.x = add(acc.x, mul(cf, v.x));
		acc.y = add(acc.y, mul(cf, v.y));
	}
	const double2 r = block_sum2(acc);
	if (threadIdx.x == 0)
		partial[blockIdx.x] = r;
}

__global__ __launch_bounds__(kBlock) void k_norm2(
	const double2 *__restrict__ a, uint64_t n, double2 *__restrict__ partial)
{
	const uint64_t stride = (uint64_t)gridDim.x * kBlock;
	double2 acc = make_double2(0.0, 0.0);
	for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
		i += stride) {
		const double2 v = a[i];
		acc.x = add(acc.x, add(mul(v.x, v.x), mul(v.y, v.y)));
	}
	const double2 r = block_sum2(acc);
	if (threadIdx.x == 0)
		partial[blockIdx.x] = r;
}

__global__ __launch_bounds__(kBlock) void k_reduce2(
	const double2 *__restrict__ partial, uint32_t n, double2 *__restrict__ out)
{
	double2 acc = make_double2(0.0, 0.0);
	for (uint32_t i = threadIdx.x; i < n; i += kBlock)
		acc = make_double2(add(acc.x, partial[i].x),
			add(acc.y, partial[i].y));
	const double2 r = block_sum2(acc);
	if (threadIdx.x == 0)
		*out = r;
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------

constexpr uint32_t kRedBlocks = 1024;

// -- per-launch event timing ------------------------------------------------

constexpr int kProfPool = 4096;
bool g_prof = false;
hipEvent_t g_pev[kProfPool][2];
int g_pkind[kProfPool];
uint32_t g_pterms[kProfPool];
uint64_t g_pamps[kProfPool];
int g_pn = 0, g_pev_made = 0;
ph2hip_prof g_pacc = {};
ph2hip_launch_rec g_pmeta[kProfPool];   // shape of each pending launch
ph2hip_launch_rec *g_rec = nullptr;     // ring of finished records
size_t g_rec_n = 0, g_rec_head = 0;
hipEvent_t g_marks[8];
bool g_marks_made = false;

int prof_drain()
{
	for (int i = 0; i < g_pn; i++) {
		float ms = 0.f;
		HIPCHK(hipEventSynchronize(g_pev[i][1]));
		HIPCHK(hipEventElapsedTime(&ms, g_pev[i][0], g_pev[i][1]));
		g_pacc.launches[g_pkind[i]]++;
		g_pacc.ms[g_pkind[i]] += ms;
		g_pacc.terms[g_pkind[i]] += g_pterms[i];
		g_pacc.amps[g_pkind[i]] += g_pamps[i];
		if (!g_rec) {
			g_rec = (ph2hip_launch_rec *)calloc(PH2HIP_REC_MAX,
				sizeof *g_rec);
			if (!g_rec)
				return fail("prof: out of host memory");
		}
		ph2hip_launch_rec r = g_pmeta[i];
		r.ms = ms;
		g_rec[(g_rec_head + g_rec_n) % PH2HIP_REC_MAX] = r;
		if (g_rec_n < PH2HIP_REC_MAX)
			g_rec_n++;
		else
			g_rec_head = (g_rec_head + 1) % PH2HIP_REC_MAX;
	}
	g_pn = 0;
	return 0;
}

// Event pair for the next launch (nullptr events when profiling is off).
// amps: amplitudes the launch reads and writes once each.
int prof_begin(int kind, uint32_t nterms, uint64_t amps, hipEvent_t *stop,
	const ph2hip_launch_rec *meta = nullptr)
{
	*stop = nullptr;
	if (!g_prof)
		return 0;
	if (g_pn == kProfPool && prof_drain())
		return -1;
	if (g_pn == g_pev_made) {
		HIPCHK(hipEventCreate(&g_pev[g_pn][0]));
		HIPCHK(hipEventCreate(&g_pev[g_pn][1]));
		g_pev_made++;
	}
	g_pkind[g_pn] = kind;
	g_pterms[g_pn] = nterms;
	g_pamps[g_pn] = amps;
	if (meta)
		g_pmeta[g_pn] = *meta;
	else
		g_pmeta[g_pn] = ph2hip_launch_rec{};
	g_pmeta[g_pn].kind = (uint32_t)kind;
	g_pmeta[g_pn].nterms = nterms;
	g_pmeta[g_pn].amps = amps;
	HIPCHK(hipEventRecord(g_pev[g_pn][0], g_stream));
	*stop = g_pev[g_pn][1];
	g_pn++;
	return 0;
}

int prof_end(hipEvent_t stop)
{
	if (stop)
		HIPCHK(hipEventRecord(stop, g_stream));
	return 0;
}

int ensure_red()
{
	if (g_red)
		return 0;
	HIPCHK(hipMalloc(&g_red, sizeof(double2) * (kRedBlocks + 1)));
	g_red_cap = kRedBlocks + 1;
	return 0;
}

uint32_t grid_for(uint64_t n, uint32_t per_block, uint32_t cap)
{
	uint64_t g = (n + per_block - 1) / per_block;
	if (g == 0)
		g = 1;
	return (uint32_t)(g < cap ? g : cap);
}

int top_bit(uint64_t v)
{
	return 63 - __builtin_clzll(v);
}

// Reduced basis of span{x & ~lowmask}: vectors with distinct leading bits,
// each zero at every other leading bit; returns the rank.
uint32_t reduce_basis(const struct ph2hip_rot *rot, size_t n, uint64_t lowmask,
	uint64_t *w, uint32_t cap)
{
	uint32_t g = 0;
	for (size_t k = 0; k < n; k++) {
		uint64_t v = rot[k].x & ~lowmask;
		for (uint32_t b = 0; b < g && v; b++)
			if (v >> top_bit(w[b]) & 1)
				v ^= w[b];
		if (!v)
			continue;
		if (g == cap)
			return cap + 1;
		// clear the new leading bit from the others, insert
		const int p = top_bit(v);
		for (uint32_t b = 0; b < g; b++)
			if (w[b] >> p & 1)
				w[b] ^= v;
		w[g++] = v;
	}
	// sort by leading bit, ascending
	for (uint32_t i = 1; i < g; i++)
		for (uint32_t j = i; j > 0 && top_bit(w[j - 1]) > top_bit(w[j]); j--) {
			const uint64_t t = w[j];
			w[j] = w[j - 1];
			w[j - 1] = t;
		}
	return g;
}

// Micro-group depth per tile size: one coset of 2^D per thread per group
// (2^(TB-D) = 256 cosets for TB <= 12; 512 for TB = 13).
constexpr int group_depth(int tb)
{
	return tb >= 12 ? 4 : tb - 8;
}

// Tile widths with a half-tile form (one coset per thread: 256 threads at
// tile 12, 512 at tile 13).  Tile 12 gains from it because the whole-tile
// LDS (64 KiB) allows two workgroups per CU and the half form four; tile 13
// goes from one 256-thread workgroup per CU to two of 512.  Narrower tiles
// already reach the register-limited count.  Batched launches too.
constexpr uint32_t kHalfTiles = PH2HIP_HALF_TILES;
constexpr uint32_t kBatchHalfTiles = (1u << 12) | (1u << 13);

bool g_pipe_init = false;
uint32_t g_pipe_mask = 0;   // bit TB set: persistent pipelined form
// Persistent half-tile form with prefetch (k_rot_fused_hpipe), per tile
// width: amplitudes of each coset prefetched (0: off) and workgroups per
// compute unit.
uint32_t g_hpipe_pf[PH2HIP_TILE_MAX + 1] = {};
uint32_t g_hpipe_wpc[PH2HIP_TILE_MAX + 1] = {};
bool g_direct = true;       // first/last groups straight to and from HBM

// Kernel forms from the environment, once: PHASE2_HIP_PIPE (mask of tile
// widths in the pipelined whole-tile form), PHASE2_HIP_PREFETCH12 and
// PHASE2_HIP_PREFETCH13 ("amplitudes:workgroups per CU" of the prefetching
// half-tile form, amplitudes 8 or 16).
void pipe_init()
{
	if (g_pipe_init)
		return;
	const char *e = getenv("PHASE2_HIP_PIPE");
	g_pipe_mask = e ? (uint32_t)strtoul(e, nullptr, 0) : 0;
	for (uint32_t tb = 12; tb <= 13; tb++) {
		char name[40];
		snprintf(name, sizeof name, "PHASE2_HIP_PREFETCH%u", tb);
		const char *f = getenv(name);
		unsigned pf = 0, wpc = 0;
		if (f && sscanf(f, "%u:%u", &pf, &wpc) == 2 &&
		    (pf == 8 || pf == 16) && wpc >= 1 && wpc <= 16) {
			g_hpipe_pf[tb] = pf;
			g_hpipe_wpc[tb] = wpc;
		}
	}
	g_pipe_init = true;
}

template <int TB>
int launch_fused(double2 *a, uint32_t nqb, const FusedPass &P,
	uint64_t xunion, uint64_t zunion)
{
	pipe_init();
	const uint64_t ntiles = UINT64_C(1) << (nqb - TB);
	hipEvent_t stop;
	ph2hip_launch_rec meta{};
	meta.nqb = nqb;
	meta.tile_bits = TB;
	meta.low_bits = P.L;
	meta.gdim = P.g;
	meta.ngroups = P.ngroups;
	meta.direct = P.dflags;
	for (uint32_t b = 0; b < P.g; b++)
		meta.pivots |= UINT64_C(1) << P.piv[b];
	meta.xunion = xunion;
	meta.zunion = zunion;
	if (prof_begin(2, P.nterms, UINT64_C(1) << nqb, &stop, &meta))
		return -1;
	if (g_pipe_mask >> TB & 1) {
		const uint64_t grid = ntiles < g_persist_wgs ? ntiles
			: g_persist_wgs;
		hipLaunchKernelGGL((k_rot_fused_pipe<TB, group_depth(TB)>),
			dim3((uint32_t)grid), dim3(kBlock), 0, g_stream, a, P,
			ntiles);
	} else {
		const uint64_t gx = ntiles < kMaxGrid ? ntiles : kMaxGrid;
		const uint64_t gy = (ntiles + gx - 1) / gx;
		if constexpr ((kHalfTiles >> TB & 1) != 0) {
			const uint32_t pf = g_hpipe_pf[TB];
			uint64_t pg = (uint64_t)g_hpipe_wpc[TB] * g_ncu;
			pg = pg < ntiles ? pg : ntiles;
			if (pg >= 8)
				pg &= ~UINT64_C(7);   // whole rounds over the XCDs
			if ((P.dflags & 4u) && pf && pg) {
				constexpr int D = group_depth(TB);
				const dim3 blk((1u << TB) >> D);
				if (pf == 16)
					hipLaunchKernelGGL((k_rot_fused_hpipe<TB, D, 16>),
						dim3((uint32_t)pg), blk, 0, g_stream, a, P,
						ntiles);
				else
					hipLaunchKernelGGL((k_rot_fused_hpipe<TB, D, 8>),
						dim3((uint32_t)pg), blk, 0, g_stream, a, P,
						ntiles);
				HIPCHK(hipGetLastError());
				return prof_end(stop);
			}
			if (P.dflags & 4u) {
				hipLaunchKernelGGL(
					(k_rot_fused_half<TB, group_depth(TB)>),
					dim3((uint32_t)gx, (uint32_t)gy),
					dim3((1u << TB) >> group_depth(TB)), 0,
					g_stream, a, P, ntiles);
				HIPCHK(hipGetLastError());
				return prof_end(stop);
			}
		}
		hipLaunchKernelGGL((k_rot_fused<TB, group_depth(TB)>),
			dim3((uint32_t)gx, (uint32_t)gy), dim3(kBlock), 0, g_stream,
			a, P, ntiles);
	}
	HIPCHK(hipGetLastError());
	return prof_end(stop);
}

// Echelon basis of span{v} for tile-coordinate vectors (leading-bit pivots,
// each basis vector zero at the other pivots); returns the rank or cap+1.
uint32_t reduce32(const uint32_t *v, size_t n, uint32_t *u, uint32_t cap)
{
	uint32_t d = 0;
	for (size_t k = 0; k < n; k++) {
		uint32_t x = v[k];
		for (uint32_t i = 0; i < d && x; i++)
			if (x >> (31 - __builtin_clz(u[i])) & 1)
				x ^= u[i];
		if (!x)
			continue;
		if (d == cap)
			return cap + 1;
		const int p = 31 - __builtin_clz(x);
		for (uint32_t i = 0; i < d; i++)
			if (u[i] >> p & 1)
				u[i] ^= x;
		u[d++] = x;
	}
	return d;
}

int g_plain_form = 2;   // unroll 2, 4 or 8; +16: non-temporal stores

template <int UNR, bool NT>
int plain_one(double2 *a, uint32_t nqb, const struct ph2hip_rot &r)
{
	const uint64_t namp = UINT64_C(1) << nqb;
	const uint64_t x = r.x, z = r.z;
	if (x == 0) {
		hipLaunchKernelGGL((k_rot_diag<UNR, NT>),
			dim3(grid_for(namp, kBlock * UNR, kMaxGrid)), dim3(kBlock),
			0, g_stream, a, namp, z, r.c, r.s);
	} else {
		const uint64_t npairs = namp / 2;
		const uint32_t y = __builtin_popcountll(x & z) & 3;
		hipLaunchKernelGGL((k_rot_pair<UNR, NT>),
			dim3(grid_for(npairs, kBlock * UNR, kMaxGrid)),
			dim3(kBlock), 0, g_stream, a, npairs, x, z,
			(uint32_t)top_bit(x), y, r.c, r.s);
	}
	HIPCHK(hipGetLastError());
	return 0;
}

// Descriptor of one fused pass over terms r[0..len) (global term index k0
// for messages) in tiles of 2^tb amplitudes.  0, or -1 if the pass does not
// fit its tile; with half set, 1 if the pass has no half-tile form (some
// exchange between consecutive groups has no split bit).
// Rank over GF(2) of v[0..n) masked by m.
uint32_t gf2_rank(const uint32_t *v, uint32_t n, uint32_t m)
{
	uint32_t b[32], r = 0;
	for (uint32_t k = 0; k < n; k++) {
		uint32_t x = v[k] & m;
		for (uint32_t i = 0; i < r && x; i++)
			if (x >> (31 - __builtin_clz(b[i])) & 1)
				x ^= b[i];
		if (!x)
			continue;
		const int p = 31 - __builtin_clz(x);
		for (uint32_t i = 0; i < r; i++)
			if (b[i] >> p & 1)
				b[i] ^= x;
		b[r++] = x;
	}
	return r;
}

// Bank vector (bits 0..3 of the 16-byte slot) of tile bit t in the
// half-tile buffer: half_slot's fold of bit t (bits 0..3 map to
// themselves).
uint32_t half_bank(uint32_t t)
{
	const uint32_t v = 1u << t;
	return (v ^ (v >> 4) ^ (v >> 5) ^ (v >> 7)) & 15u;
}

int build_pass_mode(FusedPass &P, const struct ph2hip_rot *r, size_t len,
	uint32_t tb, uint32_t low_bits, size_t k0, bool half)
{
	memset(&P, 0, sizeof P);
	// Grow the contiguous low part L while the gathered rank g fits:
	// L = tb - g(L) at the fixed point (g is non-increasing in L).
	uint32_t L = low_bits < tb ? low_bits : tb;
	uint32_t g = 0;
	uint64_t w[PH2HIP_TILE_MAX + 1];
	for (;;) {
		const uint64_t lm = (UINT64_C(1) << L) - 1;
		g = reduce_basis(r, len, lm, w, tb);
		if (L + g > tb) {
			return fail("rotate_fused: pass at term %zu needs %u+%u > %u "
			    "tile bits", k0, L, g, tb);
		}
		if (L + g == tb)
			break;
		L = tb - g;
	}
	P.L = (uint8_t)L;
	P.g = (uint8_t)g;
	P.nterms = (uint8_t)len;
	for (uint32_t b = 0; b < g; b++) {
		P.w[b] = w[b];
		P.piv[b] = (uint8_t)top_bit(w[b]);
	}
	// row part of element tid + m*kBlock that is uniform in m
	for (uint32_t m = 0; m < (1u << tb) / kBlock; m++) {
		const uint32_t rm = (m * kBlock) >> L;
		uint64_t wm = 0;
		for (uint32_t b = 0; b < g; b++)
			if (rm >> b & 1)
				wm ^= w[b];
		P.roff[m] = wm;
	}
	const uint64_t lm = (UINT64_C(1) << L) - 1;
	uint32_t xt[PH2HIP_PASS_MAX];
	for (size_t k = 0; k < len; k++) {
		const uint64_t x = r[k].x, z = r[k].z;
		uint32_t xk = (uint32_t)(x & lm), zk = (uint32_t)(z & lm);
		uint64_t rest = x & ~lm;
		for (uint32_t b = 0; b < g; b++) {
			if (x >> P.piv[b] & 1) {
				xk |= 1u << (L + b);
				rest ^= w[b];
			}
			if (__builtin_popcountll(w[b] & z) & 1)
				zk |= 1u << (L + b);
		}
		if (rest) {
			return fail("rotate_fused: term %zu outside its tile",
				k0 + k);
		}
		xt[k] = xk;
		FusedTerm &T = P.t[k];
		T.z = z;
		T.zt = zk;
		T.c = r[k].c;
		T.s = r[k].s;
		T.y = (uint8_t)(__builtin_popcountll(x & z) & 3);
	}
	// micro-groups: consecutive terms whose xt span <= D dims
	const uint32_t D = (uint32_t)group_depth((int)tb);
	uint32_t ng = 0;
	uint32_t gk[PH2HIP_PASS_MAX + 1];   // group boundaries
	uint32_t sup[PH2HIP_PASS_MAX];      // tile bits touched by each span
	for (size_t k = 0; k < len;) {
		size_t e = k + 1;
		uint32_t u[8];
		while (e < len && reduce32(xt + k, e + 1 - k, u, D) <= D)
			e++;
		const uint32_t d = reduce32(xt + k, e - k, u, D);
		uint32_t sp = 0;
		for (uint32_t m = 0; m < (1u << d); m++) {
			uint32_t v = 0;
			for (uint32_t i = 0; i < d; i++)
				if (m >> i & 1)
					v ^= u[i];
			sp |= v;
		}
		gk[ng] = (uint32_t)k;
		sup[ng++] = sp;
		k = e;
	}
	gk[ng] = (uint32_t)len;
	// Half-tile exchanges: between groups q and q+1 the tile moves through
	// LDS in two halves split by tile bit xh[q], which no vector of either
	// span touches, so that every coset lies in one half; coset-index bit
	// xk[q] (one of the top two, which select whole waves: 6/7 at tile 12,
	// 7/8 at tile 13) maps to xh[q] in both groups, so that a thread
	// writes and reads in the same half.
	uint8_t xh[PH2HIP_PASS_MAX] = {}, xk[PH2HIP_PASS_MAX] = {};
	if (half) {
		for (uint32_t q = 0; q + 1 < ng; q++) {
			const uint32_t fr = ((1u << tb) - 1) & ~(sup[q] | sup[q + 1]);
			if (!fr)
				return 1;   // no split bit: not a half-tile pass
			uint32_t h;
			if (q > 0 && (fr >> xh[q - 1] & 1)) {
				h = xh[q - 1];
			} else {
				// highest free bit above the lane bits 0..3
				const uint32_t hi = fr & ~15u;
				h = 31 - __builtin_clz(hi ? hi : fr);
			}
			// the two top coset-index bits select waves
			const uint32_t ktop = tb - D - 1;
			xh[q] = (uint8_t)h;
			xk[q] = (uint8_t)(q == 0 || h == xh[q - 1] ?
				(q ? xk[q - 1] : ktop) : 2 * ktop - 1 - xk[q - 1]);
		}
	}
	for (uint32_t gi = 0; gi < ng; gi++) {
		const size_t k = gk[gi], e = gk[gi + 1];
		// tile bits and coset-index bits reserved for the exchanges in
		// and out of this group
		uint32_t rbits = 0;
		int rpos[2] = { -1, -1 }, rbit[2] = { -1, -1 };
		if (half && gi > 0) {
			rbits |= 1u << xh[gi - 1];
			rpos[0] = xk[gi - 1];
			rbit[0] = xh[gi - 1];
		}
		if (half && gi + 1 < ng) {
			rbits |= 1u << xh[gi];
			rpos[1] = xk[gi];
			rbit[1] = xh[gi];
		}
		uint32_t u[8];
		uint32_t d = reduce32(xt + k, e - k, u, D);
		// pad with unit vectors of tile bits outside the span, highest
		// first, so that tile bits 0..3 stay free for the lanes where
		// possible (whole 256-byte runs per quarter-wave)
		for (uint32_t bit = tb; d < D && bit-- > 0;) {
			if (rbits >> bit & 1)
				continue;
			uint32_t x = 1u << bit;
			for (uint32_t i = 0; i < d && x; i++)
				if (x >> (31 - __builtin_clz(u[i])) & 1)
					x ^= u[i];
			if (!x)
				continue;
			const int p = 31 - __builtin_clz(x);
			for (uint32_t i = 0; i < d; i++)
				if (u[i] >> p & 1)
					u[i] ^= x;
			u[d++] = x;
		}
		if (d < D)
			return half ? 1 : fail("rotate_fused: group padding");
		for (uint32_t i = 1; i < d; i++)   // ascending pivots
			for (uint32_t j = i; j > 0 && __builtin_clz(u[j - 1]) <
				__builtin_clz(u[j]); j--) {
				const uint32_t t = u[j];
				u[j] = u[j - 1];
				u[j - 1] = t;
			}
		FusedGroup &G = P.grp[gi];
		memset(&G, 0, sizeof G);
		for (uint32_t i = 0; i < D; i++) {
			G.u[i] = (uint16_t)u[i];
			G.piv[i] = (uint8_t)(31 - __builtin_clz(u[i]));
		}
		// Coset index bits -> free tile bits.  The first four (lanes
		// 0..15 of a wave) take free bits whose bank vectors are
		// independent.  In the whole-tile form tile bit t < 12 lands
		// on bank bit t mod 4, one bit per residue.  In the half-tile
		// form the bank vector is half_bank(t): the first three lanes
		// (served together by a 16-byte write) take bits independent
		// mod 8, the fourth one independent mod 16 (a 16-byte read
		// serves 16 lanes), in ascending order: tile bits 0..3 first
		// where free, whole 256-byte runs for the groups that move
		// their cosets to and from HBM; tile bit tb - 1, which the
		// exchange moves to the split bit, last.
		// Reserved index bits take their exchange split bits.
		{
			uint32_t used = rbits, npos = 0;
			for (uint32_t i = 0; i < D; i++)
				used |= 1u << G.piv[i];
			bool taken[16] = {};
			for (int r = 0; r < 2; r++)
				if (rpos[r] >= 0) {
					G.jpos[rpos[r]] = (uint8_t)rbit[r];
					taken[rpos[r]] = true;
				}
			if (half) {
				// round 0: three lanes independent mod 8;
				// round 1: a fourth independent mod 16;
				// round 2: any free bits if those do not exist
				uint32_t bv[5], nl = 0;
				for (int round = 0; round < 3 && nl < 4; round++)
					for (uint32_t tt = 0; tt < tb &&
						nl < (round == 0 ? 3u : 4u); tt++) {
						if (used >> tt & 1)
							continue;
						bv[nl] = half_bank(tt);
						const uint32_t m = round == 0 ? 7u : 15u;
						if (round < 2 &&
						    gf2_rank(bv, nl + 1, m) != nl + 1)
							continue;
						used |= 1u << tt;
						while (taken[npos])
							npos++;
						G.jpos[npos++] = (uint8_t)tt;
						nl++;
					}
			} else {
				for (uint32_t res = 0; res < 4; res++)
					for (uint32_t t = res; t < tb && t < 12; t += 4)
						if (!(used >> t & 1)) {
							used |= 1u << t;
							while (taken[npos])
								npos++;
							G.jpos[npos++] = (uint8_t)t;
							break;
						}
			}
			for (uint32_t t = 0; t < tb; t++)
				if (!(used >> t & 1)) {
					used |= 1u << t;
					while (taken[npos])
						npos++;
					G.jpos[npos++] = (uint8_t)t;
				}
			while (npos < 16 && taken[npos])
				npos++;
			if (npos != tb - D || used != (1u << tb) - 1)
				return fail("rotate_fused: coset map");
		}
		G.first = (uint8_t)k;
		G.count = (uint8_t)(e - k);
		G.xh = xh[gi];
		G.xk = xk[gi];
		for (size_t q = k; q < e; q++) {
			uint32_t xp = 0, zp = 0, rest = xt[q];
			for (uint32_t i = 0; i < D; i++) {
				if (xt[q] >> G.piv[i] & 1) {
					xp |= 1u << i;
					rest ^= u[i];
				}
				if (__builtin_popcount(u[i] & P.t[q].zt) & 1)
					zp |= 1u << i;
			}
			if (rest)
				return fail("rotate_fused: group basis");
			P.t[q].xtp = (uint8_t)xp;
			P.t[q].ztp = (uint8_t)zp;
			// bit w of the top half: parity(w & ztp), the sign of
			// coset register w's partner term (tile coordinates use
			// at most the low 13 bits of zt)
			uint32_t sgn = 0;
			for (uint32_t w = 0; w < (1u << D); w++)
				sgn |= (uint32_t)(__builtin_popcount(w & zp) & 1) <<
					w;
			P.t[q].zt = (P.t[q].zt & 0xffffu) | (sgn << 16);
		}
	}
	P.ngroups = (uint8_t)ng;
	// First and last group straight between HBM and registers when their
	// coset-index bits 0..3 are tile bits 0..3 inside rows of >= 16; for a
	// half-tile pass always (the tile is never staged whole in LDS).
	if ((g_direct && L >= 4) || half) {
		for (int side = 0; side < 2; side++) {
			const FusedGroup &G = P.grp[side ? ng - 1 : 0];
			if (!half && (G.jpos[0] != 0 || G.jpos[1] != 1 ||
			    G.jpos[2] != 2 || G.jpos[3] != 3))
				continue;
			for (uint32_t e = 0; e < (1u << D); e++) {
				uint32_t cw = 0;
				for (uint32_t i = 0; i < D; i++)
					if (e >> i & 1)
						cw ^= G.u[i];
				uint64_t off = cw & lm;
				for (uint32_t b = 0; b < g; b++)
					if ((cw >> (L + b)) & 1)
						off ^= w[b];
				P.dcw[side][e] = off;
			}
			P.dflags |= 1u << side;
		}
	}
	if (half)
		P.dflags |= 4u;
	return 0;
}

// Bit tb set: passes of tile width tb take the half-tile form where they
// have one (k_rot_fused_half), else the whole-tile form.  Default: every
// width that has the form (measured at 33 qubits, tile 12: 53.9 ms per
// pass against 61.6 ms for the same greedy passes in the whole-tile form).
uint32_t g_half_mask = kHalfTiles;
bool g_half_init = false;

uint32_t half_mask()
{
	if (!g_half_init) {
		const char *e = getenv("PHASE2_HIP_HALF");
		if (e)
			g_half_mask = (uint32_t)strtoul(e, nullptr, 0) &
				kHalfTiles;
		g_half_init = true;
	}
	return g_half_mask;
}

// A single-register pass in the form its launch will take.
int build_pass(FusedPass &P, const struct ph2hip_rot *r, size_t len,
	uint32_t tb, uint32_t low_bits, size_t k0)
{
	if (half_mask() >> tb & 1) {
		const int rc = build_pass_mode(P, r, len, tb, low_bits, k0,
			true);
		if (rc <= 0)
			return rc;
	}
	return build_pass_mode(P, r, len, tb, low_bits, k0, false);
}

// Grow-only device buffer.
int dev_reserve(void **buf, size_t *cap, size_t bytes)
{
	if (bytes <= *cap)
		return 0;
	if (*buf) {
		HIPCHK(hipStreamSynchronize(g_stream));
		HIPCHK(hipFree(*buf));
		*buf = nullptr;
		*cap = 0;
	}
	HIPCHK(hipMalloc(buf, bytes));
	*cap = bytes;
	return 0;
}

void *g_bdesc = nullptr, *g_bitem = nullptr;
size_t g_bdesc_cap = 0, g_bitem_cap = 0;

template <int TB>
int launch_batch(double2 *amp, const FusedPass *descs, const BatchItem *items,
	uint32_t nitems, uint64_t ntiles, bool half)
{
	const uint64_t gx = ntiles < kMaxGrid ? ntiles : kMaxGrid;
	const uint64_t gy = (ntiles + gx - 1) / gx;
	if constexpr ((kBatchHalfTiles >> TB & 1) != 0) {
		if (half) {
			hipLaunchKernelGGL(
				(k_rot_fused_batch<TB, group_depth(TB), true>),
				dim3((uint32_t)gx, (uint32_t)gy),
				dim3((1u << TB) >> group_depth(TB)), 0, g_stream,
				amp, descs, items, nitems, ntiles);
			HIPCHK(hipGetLastError());
			return 0;
		}
	}
	hipLaunchKernelGGL((k_rot_fused_batch<TB, group_depth(TB)>),
		dim3((uint32_t)gx, (uint32_t)gy), dim3(kBlock), 0, g_stream, amp,
		descs, items, nitems, ntiles);
	HIPCHK(hipGetLastError());
	return 0;
}

// Replay of repeated fused sweeps on narrow registers.  A Trotter step
// applies the same circ_cache block (same terms, angles and schedule) to
// the same register every step; below ~20 qubits its passes take a few
// microseconds each and launching them one by one costs about as much.
// The first time a call is seen it is remembered (a copy of its
// arguments); the second time its launches are captured into a hipGraph
// (kernel nodes only: the pass descriptors are kernel arguments); from
// then on the graph is replayed.  Calls match only if every argument is
// equal, compared in full, so a replay launches exactly what the direct
// path would.
constexpr int kReplaySlots = 32;

struct ReplaySlot {
	uint64_t key, used;
	void *amp;
	uint32_t nqb;
	size_t n, npass;
	struct ph2hip_rot *rot;
	uint32_t *sched;            // pass_len, tile_bits, low_bits
	hipGraph_t graph;
	hipGraphExec_t exec;
};

ReplaySlot g_replay[kReplaySlots];
uint64_t g_replay_clock = 0, g_replay_hits = 0;
bool g_replay_init = false;
uint32_t g_replay_max_qb = 20;

uint64_t fnv_bytes(uint64_t h, const void *p, size_t n)
{
	const unsigned char *b = static_cast<const unsigned char *>(p);
	for (size_t i = 0; i < n; i++)
		h = (h ^ b[i]) * UINT64_C(0x100000001b3);
	return h;
}

void replay_clear(ReplaySlot &r)
{
	if (r.exec)
		(void)hipGraphExecDestroy(r.exec);
	if (r.graph)
		(void)hipGraphDestroy(r.graph);
	free(r.rot);
	free(r.sched);
	r = ReplaySlot{};
}

bool replay_same(const ReplaySlot &r, uint64_t key, void *amp, uint32_t nqb,
	const struct ph2hip_rot *rot, size_t n, const uint32_t *const sch[3],
	size_t npass)
{
	if (!r.rot || r.key != key || r.amp != amp || r.nqb != nqb ||
	    r.n != n || r.npass != npass ||
	    memcmp(r.rot, rot, sizeof *rot * n) != 0)
		return false;
	for (int a = 0; a < 3; a++)
		if (memcmp(r.sched + a * npass, sch[a],
			    sizeof(uint32_t) * npass) != 0)
			return false;
	return true;
}

// The slot holding this call, or a fresh one (least recently used)
// remembering it; *seen tells which.  nullptr if out of host memory.
ReplaySlot *replay_slot(void *amp, uint32_t nqb, const struct ph2hip_rot *rot,
	size_t n, const uint32_t *const sch[3], size_t npass, bool *seen)
{
	uint64_t key = fnv_bytes(UINT64_C(0xcbf29ce484222325), rot,
		sizeof *rot * n);
	for (int a = 0; a < 3; a++)
		key = fnv_bytes(key, sch[a], sizeof(uint32_t) * npass);
	ReplaySlot *old = &g_replay[0];
	for (ReplaySlot &r : g_replay) {
		if (replay_same(r, key, amp, nqb, rot, n, sch, npass)) {
			r.used = ++g_replay_clock;
			*seen = true;
			return &r;
		}
		if (r.used < old->used)
			old = &r;
	}
	*seen = false;
	replay_clear(*old);
	old->rot = static_cast<struct ph2hip_rot *>(malloc(sizeof *rot * n));
	old->sched = static_cast<uint32_t *>(malloc(sizeof(uint32_t) * 3 *
		npass));
	if (!old->rot || !old->sched) {
		replay_clear(*old);
		return nullptr;
	}
	memcpy(old->rot, rot, sizeof *rot * n);
	for (int a = 0; a < 3; a++)
		memcpy(old->sched + a * npass, sch[a], sizeof(uint32_t) * npass);
	old->key = key;
	old->amp = amp;
	old->nqb = nqb;
	old->n = n;
	old->npass = npass;
	old->used = ++g_replay_clock;
	return old;
}

} // namespace

struct ph2hip_batch_plan {
	FusedPass *desc;      // descriptors, job-major
	BatchItem *items;     // work items, launch-major
	size_t ndesc, maxp;
	size_t nlaunch;       // one or two per pass index (half-tile, whole)
	size_t *lstart;       // launch l: items [lstart[l], lstart[l+1])
	uint64_t *ltiles;     // tiles of launch l
	uint8_t *lhalf;       // launch l in the half-tile form
	uint32_t *lpass;      // pass index of launch l
	uint32_t tile_bits;
	uint64_t namp_total;
};

namespace {

// The batched rounds' graph, built once and replayed: two kernel nodes
// per pass index (the half-tile form, then the whole-tile form; one node
// at tile widths without a half-tile form), chained in that order.  Each
// round sets every node's arguments (its work items, or none: a node
// without work exits at once) and launches the graph; it is rebuilt only
// when a round needs more pass indices or another tile width.  The node
// arguments change only after the previous replay has finished (an event
// recorded behind it).
struct RoundGraph {
	uint32_t tile_bits = 0, per = 0;
	size_t npass = 0;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	hipGraphNode_t *nodes = nullptr;
	hipEvent_t done = nullptr;
	bool pending = false;
};
RoundGraph g_rg;
uint64_t g_rg_builds = 0, g_rg_replays = 0;

void rg_clear()
{
	if (g_rg.pending && g_rg.done)
		(void)hipEventSynchronize(g_rg.done);
	if (g_rg.exec)
		(void)hipGraphExecDestroy(g_rg.exec);
	if (g_rg.graph)
		(void)hipGraphDestroy(g_rg.graph);
	if (g_rg.done)
		(void)hipEventDestroy(g_rg.done);
	free(g_rg.nodes);
	g_rg = RoundGraph{};
}

// Arguments of one batched node (k_rot_fused_batch's parameters).
struct BatchArgs {
	double2 *amp;
	const FusedPass *descs;
	const BatchItem *items;
	uint32_t nitems;
	uint64_t ntiles;
	void *ptr[5];
};

template <int TB>
void batch_node(hipKernelNodeParams &kp, BatchArgs &ba, bool half)
{
	constexpr int D = group_depth(TB);
	ba.ptr[0] = &ba.amp;
	ba.ptr[1] = &ba.descs;
	ba.ptr[2] = &ba.items;
	ba.ptr[3] = &ba.nitems;
	ba.ptr[4] = &ba.ntiles;
	const uint64_t nt = ba.ntiles;
	const uint64_t gx = nt == 0 ? 1 : nt < kMaxGrid ? nt : kMaxGrid;
	const uint64_t gy = nt == 0 ? 1 : (nt + gx - 1) / gx;
	kp = hipKernelNodeParams{};
	if constexpr ((kBatchHalfTiles >> TB & 1) != 0) {
		if (half) {
			kp.func = reinterpret_cast<void *>(
				&k_rot_fused_batch<TB, D, true>);
			kp.blockDim = dim3((1u << TB) >> D);
		}
	}
	if (!kp.func) {
		kp.func = reinterpret_cast<void *>(&k_rot_fused_batch<TB, D>);
		kp.blockDim = dim3(kBlock);
	}
	kp.gridDim = dim3((uint32_t)gx, (uint32_t)gy);
	kp.sharedMemBytes = 0;
	kp.kernelParams = ba.ptr;
	kp.extra = nullptr;
}

void batch_node_tb(uint32_t tb, hipKernelNodeParams &kp, BatchArgs &ba,
	bool half)
{
	switch (tb) {
	case 9: batch_node<9>(kp, ba, half); break;
	case 10: batch_node<10>(kp, ba, half); break;
	case 11: batch_node<11>(kp, ba, half); break;
	case 12: batch_node<12>(kp, ba, half); break;
	default: batch_node<13>(kp, ba, half); break;
	}
}

int rg_build(uint32_t tb, size_t npass, double2 *amp)
{
	rg_clear();
	g_rg.tile_bits = tb;
	g_rg.per = (kBatchHalfTiles >> tb & 1) ? 2 : 1;
	g_rg.npass = npass;
	const size_t nn = npass * g_rg.per;
	g_rg.nodes = static_cast<hipGraphNode_t *>(calloc(nn ? nn : 1,
		sizeof(hipGraphNode_t)));
	if (!g_rg.nodes)
		return fail("round graph: out of host memory");
	HIPCHK(hipEventCreateWithFlags(&g_rg.done, hipEventDisableTiming));
	HIPCHK(hipGraphCreate(&g_rg.graph, 0));
	for (size_t i = 0; i < nn; i++) {
		BatchArgs ba{ amp, nullptr, nullptr, 0, 0, {} };
		hipKernelNodeParams kp;
		batch_node_tb(tb, kp, ba, g_rg.per == 2 && i % 2 == 0);
		HIPCHK(hipGraphAddKernelNode(&g_rg.nodes[i], g_rg.graph,
			i ? &g_rg.nodes[i - 1] : nullptr, i ? 1 : 0, &kp));
	}
	HIPCHK(hipGraphInstantiate(&g_rg.exec, g_rg.graph, nullptr, nullptr, 0));
	g_rg_builds++;
	return 0;
}

// One round of a plan through the round graph (descriptors and items
// already uploaded to g_bdesc / g_bitem).
int rg_launch(const struct ph2hip_batch_plan *pl, double2 *amp)
{
	const uint32_t tb = pl->tile_bits;
	if (!g_rg.exec || g_rg.tile_bits != tb || g_rg.npass < pl->maxp) {
		const size_t np = g_rg.tile_bits == tb && g_rg.npass > pl->maxp ?
			g_rg.npass : pl->maxp;
		if (rg_build(tb, np, amp))
			return -1;
	} else if (g_rg.pending) {
		HIPCHK(hipEventSynchronize(g_rg.done));
	}
	g_rg.pending = false;
	// launch of each node: l for (pass p, form), or none
	const size_t nn = g_rg.npass * g_rg.per;
	long *lof = static_cast<long *>(malloc(sizeof(long) * (nn ? nn : 1)));
	if (!lof)
		return fail("round graph: out of host memory");
	for (size_t i = 0; i < nn; i++)
		lof[i] = -1;
	for (size_t l = 0; l < pl->nlaunch; l++) {
		const size_t slot = g_rg.per == 2 ?
			2 * pl->lpass[l] + (pl->lhalf[l] ? 0 : 1) : pl->lpass[l];
		lof[slot] = (long)l;
	}
	int rc = 0;
	for (size_t i = 0; i < nn && !rc; i++) {
		BatchArgs ba{ amp, static_cast<const FusedPass *>(g_bdesc),
			static_cast<const BatchItem *>(g_bitem), 0, 0, {} };
		if (lof[i] >= 0) {
			const size_t l = (size_t)lof[i];
			ba.items += pl->lstart[l];
			ba.nitems = (uint32_t)(pl->lstart[l + 1] - pl->lstart[l]);
			ba.ntiles = pl->ltiles[l];
		}
		hipKernelNodeParams kp;
		batch_node_tb(tb, kp, ba, g_rg.per == 2 && i % 2 == 0);
		if (hipGraphExecKernelNodeSetParams(g_rg.exec, g_rg.nodes[i], &kp)
			!= hipSuccess)
			rc = fail("round graph: node %zu arguments", i);
	}
	free(lof);
	if (rc)
		return rc;
	HIPCHK(hipGraphLaunch(g_rg.exec, g_stream));
	HIPCHK(hipEventRecord(g_rg.done, g_stream));
	g_rg.pending = true;
	g_rg_replays++;
	return 0;
}

} // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

int ph2hip_init(int device)
{
	if (g_stream)
		return 0;
	if (device < 0) {
		const char *env = getenv("PHASE2_HIP_DEVICE");
		device = env ? atoi(env) : 0;
	}
	int ndev = 0;
	HIPCHK(hipGetDeviceCount(&ndev));
	if (device >= ndev)
		return fail("device %d not present (%d devices)", device, ndev);
	HIPCHK(hipSetDevice(device));
	if (device < kMaxDevices && !g_scratch[device]) {
		void *p = nullptr;
		const hipError_t e = hipGetSymbolAddress(&p,
			HIP_SYMBOL(scratch));
		if (e != hipSuccess || !p)
			return fail("init: address of the gather scratch on "
				"device %d: %s", device, hipGetErrorString(e));
		g_scratch[device] = p;
	}
	HIPCHK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
	g_device = device;
	int ncu = 0;
	if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
		    device) == hipSuccess && ncu > 0) {
		g_persist_wgs = 2 * (uint64_t)ncu;
		g_ncu = (uint32_t)ncu;
	}
	const char *pw = getenv("PHASE2_HIP_FUSED_WGS");
	if (pw && atoll(pw) > 0)
		g_persist_wgs = (uint64_t)atoll(pw);
	return 0;
}

void ph2hip_shutdown(void)
{
	if (!g_stream)
		return;
	(void)hipStreamSynchronize(g_stream);
	for (ReplaySlot &r : g_replay)
		replay_clear(r);
	rg_clear();
	if (g_red)
		(void)hipFree(g_red);
	g_red = nullptr;
	(void)hipStreamDestroy(g_stream);
	g_stream = nullptr;
}

const char *ph2hip_strerror(void)
{
	return g_err;
}

int ph2hip_synchronize(void)
{
	if (ensure_init())
		return -1;
	HIPCHK(hipStreamSynchronize(g_stream));
	return 0;
}

int ph2hip_malloc(void **dptr, size_t bytes)
{
	if (ensure_init())
		return -1;
	*dptr = nullptr;
	HIPCHK(hipMalloc(dptr, bytes ? bytes : 16));
	return 0;
}

int ph2hip_free(void *dptr)
{
	if (!dptr)
		return 0;
	if (g_stream)
		HIPCHK(hipStreamSynchronize(g_stream));
	HIPCHK(hipFree(dptr));
	return 0;
}

int ph2hip_h2d(void *dst, const void *src, size_t bytes)
{
	if (ensure_init())
		return -1;
	HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_stream));
	HIPCHK(hipStreamSynchronize(g_stream));
	return 0;
}

int ph2hip_d2h(void *dst, const void *src, size_t bytes)
{
	if (ensure_init())
		return -1;
	HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_stream));
	HIPCHK(hipStreamSynchronize(g_stream));
	return 0;
}

int ph2hip_d2d(void *dst, const void *src, size_t bytes)
{
	if (ensure_init())
		return -1;
	HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice,
		g_stream));
	return 0;
}

int ph2hip_zero(void *amp, uint64_t namp)
{
	if (ensure_init())
		return -1;
	hipLaunchKernelGGL(k_fill, dim3(grid_for(namp, kBlock * 4, 1u << 20)),
		dim3(kBlock), 0, g_stream, (double2 *)amp, namp);
	HIPCHK(hipGetLastError());
	return 0;
}

int ph2hip_rotate_plain(void *amp, uint32_t nqb, const struct ph2hip_rot *rot,
	size_t n)
{
	if (ensure_init())
		return -1;
	if (nqb == 0 || nqb > 40)
		return fail("rotate_plain: nqb=%u out of range", nqb);
	double2 *a = (double2 *)amp;
	for (size_t k = 0; k < n; k++) {
		const uint64_t x = rot[k].x, z = rot[k].z;
		if (x >> nqb)
			return fail("rotate_plain: term %zu flips qubit >= %u", k,
				nqb);
		hipEvent_t stop;
		ph2hip_launch_rec meta{};
		meta.nqb = nqb;
		meta.xunion = x;
		meta.zunion = z;
		if (prof_begin(x == 0 ? 1 : 0, 1, UINT64_C(1) << nqb, &stop,
			    &meta))
			return -1;
		int rc;
		switch (g_plain_form) {
		case 4: rc = plain_one<4, false>(a, nqb, rot[k]); break;
		case 8: rc = plain_one<8, false>(a, nqb, rot[k]); break;
		case 2 + 16: rc = plain_one<2, true>(a, nqb, rot[k]); break;
		case 4 + 16: rc = plain_one<4, true>(a, nqb, rot[k]); break;
		default: rc = plain_one<2, false>(a, nqb, rot[k]); break;
		}
		if (rc || prof_end(stop))
			return -1;
	}
	return 0;
}

int ph2hip_set_plain_form(int form)
{
	g_plain_form = form;
	return 0;
}

uint32_t ph2hip_span_rank(const struct ph2hip_rot *rot, size_t n,
	uint32_t low_bits)
{
	uint64_t w[64];
	const uint64_t lowmask = low_bits >= 64 ? ~UINT64_C(0)
		: (UINT64_C(1) << low_bits) - 1;
	return reduce_basis(rot, n, lowmask, w, 64);
}

int ph2hip_rotate_fused(void *amp, uint32_t nqb, const struct ph2hip_rot *rot,
	size_t n, const uint32_t *pass_len, const uint32_t *tile_bits,
	size_t npass, const uint32_t *low_bits)
{
	if (ensure_init())
		return -1;
	if (nqb == 0 || nqb > 40)
		return fail("rotate_fused: nqb=%u out of range", nqb);
	size_t tot = 0;
	for (size_t p = 0; p < npass; p++) {
		if (pass_len[p] == 0 || pass_len[p] > PH2HIP_PASS_MAX)
			return fail("rotate_fused: pass %zu has %u terms", p,
				pass_len[p]);
		const uint32_t tb = tile_bits[p];
		if (tb < PH2HIP_TILE_MIN || tb > PH2HIP_TILE_MAX || tb > nqb)
			return fail("rotate_fused: tile_bits %u invalid for "
				    "nqb=%u", tb, nqb);
		tot += pass_len[p];
	}
	if (tot != n)
		return fail("rotate_fused: passes cover %zu of %zu terms", tot, n);
	for (size_t k = 0; k < n; k++)
		if (rot[k].x >> nqb)
			return fail("rotate_fused: term %zu flips qubit >= %u", k,
				nqb);

	ReplaySlot *slot = nullptr;
	ph2hip_replay_max_qb();
	if (nqb <= g_replay_max_qb && !g_prof && n > 0) {
		const uint32_t *const sch[3] = { pass_len, tile_bits, low_bits };
		bool seen = false;
		slot = replay_slot(amp, nqb, rot, n, sch, npass, &seen);
		if (slot && slot->exec) {
			HIPCHK(hipGraphLaunch(slot->exec, g_stream));
			g_replay_hits++;
			return 0;
		}
		if (!seen)
			slot = nullptr;   // first sight: launch directly
	}

	// Build every pass first, so nothing is applied if one does not fit.
	FusedPass *passes = (FusedPass *)calloc(npass ? npass : 1,
		sizeof(FusedPass));
	if (!passes)
		return fail("rotate_fused: out of host memory");
	size_t k0 = 0;
	for (size_t p = 0; p < npass; p++) {
		if (build_pass(passes[p], rot + k0, pass_len[p], tile_bits[p],
			    low_bits[p], k0)) {
			free(passes);
			return -1;
		}
		k0 += pass_len[p];
	}
	double2 *a = (double2 *)amp;
	int rc = 0;
	// second sight of a call on a narrow register: capture its launches
	if (slot && hipStreamBeginCapture(g_stream,
			    hipStreamCaptureModeThreadLocal) != hipSuccess)
		slot = nullptr;
	k0 = 0;
	for (size_t p = 0; p < npass && rc == 0; p++) {
		uint64_t xu = 0, zu = 0;
		for (size_t k = k0; k < k0 + pass_len[p]; k++) {
			xu |= rot[k].x;
			zu |= rot[k].z;
		}
		k0 += pass_len[p];
		const FusedPass &P = passes[p];
		switch (tile_bits[p]) {
		case 9: rc = launch_fused<9>(a, nqb, P, xu, zu); break;
		case 10: rc = launch_fused<10>(a, nqb, P, xu, zu); break;
		case 11: rc = launch_fused<11>(a, nqb, P, xu, zu); break;
		case 12: rc = launch_fused<12>(a, nqb, P, xu, zu); break;
		case 13: rc = launch_fused<13>(a, nqb, P, xu, zu); break;
		default: rc = fail("rotate_fused: tile_bits"); break;
		}
	}
	free(passes);
	if (slot) {
		hipGraph_t graph = nullptr;
		const hipError_t e = hipStreamEndCapture(g_stream, &graph);
		if (!rc && e != hipSuccess)
			rc = fail("rotate_fused: end capture: %s",
				hipGetErrorString(e));
		if (!rc && hipGraphInstantiate(&slot->exec, graph, nullptr,
			    nullptr, 0) != hipSuccess) {
			slot->exec = nullptr;
			rc = fail("rotate_fused: graph instantiate");
		}
		slot->graph = graph;
		if (!rc && hipGraphLaunch(slot->exec, g_stream) != hipSuccess)
			rc = fail("rotate_fused: graph launch");
		if (rc)
			replay_clear(*slot);
	}
	return rc;
}

uint32_t ph2hip_replay_max_qb(void)
{
	if (!g_replay_init) {
		const char *e = getenv("PHASE2_HIP_REPLAY_QB");
		if (e)
			g_replay_max_qb = (uint32_t)strtoul(e, nullptr, 0);
		g_replay_init = true;
	}
	return g_replay_max_qb;
}

int ph2hip_set_replay(uint32_t max_qb)
{
	if (ensure_init())
		return -1;
	g_replay_init = true;
	g_replay_max_qb = max_qb;
	HIPCHK(hipStreamSynchronize(g_stream));
	for (ReplaySlot &r : g_replay)
		replay_clear(r);
	return 0;
}

uint64_t ph2hip_replay_hits(void)
{
	return g_replay_hits;
}

int ph2hip_gather(const void *amp, uint32_t nqb, const uint64_t *idx,
	size_t n, double *out)
{
	if (ensure_init())
		return -1;
	if (n == 0)
		return 0;
	const uint64_t namp = UINT64_C(1) << nqb;
	for (size_t k = 0; k < n; k++)
		if (idx[k] >= namp)
			return fail("gather: index %llu beyond 2^%u",
				(unsigned long long)idx[k], nqb);
	uint64_t *didx = nullptr;
	double2 *dout = nullptr;
	HIPCHK(hipMalloc(&didx, sizeof(uint64_t) * n));
	if (hipMalloc(&dout, sizeof(double2) * n) != hipSuccess) {
		(void)hipFree(didx);
		return fail("gather: hipMalloc");
	}
	int rc = 0;
	if (hipMemcpyAsync(didx, idx, sizeof(uint64_t) * n,
		    hipMemcpyHostToDevice, g_stream) != hipSuccess)
		rc = fail("gather: copy in");
	if (!rc) {
		hipLaunchKernelGGL(k_gather, dim3(grid_for(n, kBlock, kMaxGrid)),
			dim3(kBlock), 0, g_stream, (const double2 *)amp, didx,
			dout, (uint64_t)n);
		if (hipGetLastError() != hipSuccess)
			rc = fail("gather: launch");
	}
	if (!rc && hipMemcpyAsync(out, dout, sizeof(double2) * n,
			   hipMemcpyDeviceToHost, g_stream) != hipSuccess)
		rc = fail("gather: copy out");
	if (hipStreamSynchronize(g_stream) != hipSuccess && !rc)
		rc = fail("gather: sync");
	(void)hipFree(didx);
	(void)hipFree(dout);
	return rc;
}

int ph2hip_gather_noalloc(const void *amp, uint32_t nqb, const uint64_t *idx,
	size_t n, double *out)
{
	if (ensure_init())
		return -1;
	if (n == 0)
		return 0;
	if (n > PH2HIP_GATHER_NOALLOC_MAX)
		return fail("gather_noalloc: %zu indices, at most %d", n,
			PH2HIP_GATHER_NOALLOC_MAX);
	const uint64_t namp = UINT64_C(1) << nqb;
	for (size_t k = 0; k < n; k++)
		if (idx[k] >= namp)
			return fail("gather_noalloc: index %llu beyond 2^%u",
				(unsigned long long)idx[k], nqb);
	void *p = g_device >= 0 && g_device < kMaxDevices ?
		g_scratch[g_device] : nullptr;
	if (!p)
		return fail("gather_noalloc: no scratch address for device %d",
			g_device);
	// indices first, amplitudes after the largest index block; the copy
	// in overwrites what the previous call left, the kernel every output
	uint64_t *didx = static_cast<uint64_t *>(p);
	double2 *dout = reinterpret_cast<double2 *>(
		didx + PH2HIP_GATHER_NOALLOC_MAX);
	HIPCHK(hipMemcpyAsync(didx, idx, sizeof(uint64_t) * n,
		hipMemcpyHostToDevice, g_stream));
	hipLaunchKernelGGL(k_gather, dim3(grid_for(n, kBlock, kMaxGrid)),
		dim3(kBlock), 0, g_stream, (const double2 *)amp, didx, dout,
		(uint64_t)n);
	HIPCHK(hipGetLastError());
	HIPCHK(hipMemcpyAsync(out, dout, sizeof(double2) * n,
		hipMemcpyDeviceToHost, g_stream));
	HIPCHK(hipStreamSynchronize(g_stream));
	return 0;
}

struct ph2hip_sc {
	uint64_t *ma, *mb;
	double *da, *db;
	size_t Ma, Mb;
	uint64_t or_a, or_b;  // OR of all masks (index-range check)
};

int ph2hip_sc_create(struct ph2hip_sc **out, const uint64_t *mask_a,
	size_t Ma, const uint64_t *mask_b, size_t Mb)
{
	if (ensure_init())
		return -1;
	*out = nullptr;
	struct ph2hip_sc *sc = (struct ph2hip_sc *)calloc(1, sizeof *sc);
	if (!sc)
		return fail("sc_create: out of host memory");
	sc->Ma = Ma;
	sc->Mb = Mb;
	for (size_t i = 0; i < Ma; i++)
		sc->or_a |= mask_a[i];
	for (size_t i = 0; i < Mb; i++)
		sc->or_b |= mask_b[i];
	if (hipMalloc(&sc->ma, sizeof(uint64_t) * (Ma ? Ma : 1)) != hipSuccess ||
	    hipMalloc(&sc->mb, sizeof(uint64_t) * (Mb ? Mb : 1)) != hipSuccess ||
	    hipMalloc(&sc->da, sizeof(double) * (Ma ? Ma : 1)) != hipSuccess ||
	    hipMalloc(&sc->db, sizeof(double) * (Mb ? Mb : 1)) != hipSuccess) {
		ph2hip_sc_destroy(sc);
		return fail("sc_create: hipMalloc");
	}
	if (hipMemcpy(sc->ma, mask_a, sizeof(uint64_t) * Ma,
		    hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(sc->mb, mask_b, sizeof(uint64_t) * Mb,
		    hipMemcpyHostToDevice) != hipSuccess) {
		ph2hip_sc_destroy(sc);
		return fail("sc_create: copy");
	}
	*out = sc;
	return 0;
}

int ph2hip_sc_set_dets(struct ph2hip_sc *sc, const double *det_a,
	const double *det_b)
{
	if (ensure_init())
		return -1;
	HIPCHK(hipMemcpyAsync(sc->da, det_a, sizeof(double) * sc->Ma,
		hipMemcpyHostToDevice, g_stream));
	HIPCHK(hipMemcpyAsync(sc->db, det_b, sizeof(double) * sc->Mb,
		hipMemcpyHostToDevice, g_stream));
	HIPCHK(hipStreamSynchronize(g_stream));
	return 0;
}

void ph2hip_sc_destroy(struct ph2hip_sc *sc)
{
	if (!sc)
		return;
	if (g_stream)
		(void)hipStreamSynchronize(g_stream);
	(void)hipFree(sc->ma);
	(void)hipFree(sc->mb);
	(void)hipFree(sc->da);
	(void)hipFree(sc->db);
	free(sc);
}

static int sc_check(const struct ph2hip_sc *sc, uint32_t nqb,
	uint32_t n_sites, int tapered)
{
	if (n_sites == 0 || n_sites > 32)
		return fail("sc: n_sites=%u out of range", n_sites);
	if (sc->or_b && (2 * n_sites > 64))
		return fail("sc: beta bits beyond 64");
	uint64_t bound = sc->or_a | (sc->or_b << n_sites);
	if (tapered) {
		const uint64_t lo = (bound >> 1) &
			((UINT64_C(1) << (n_sites - 1)) - 1);
		bound = lo | ((bound >> (n_sites + 1)) << (n_sites - 1));
	}
	if (nqb < 64 && (bound >> nqb))
		return fail("sc: determinant index beyond the %u-qubit register "
			    "(n_sites=%u)", nqb, n_sites);
	return 0;
}

int ph2hip_sc_scatter(void *amp, uint32_t nqb, const struct ph2hip_sc *sc,
	uint32_t n_sites, double weight, int tapered, int accumulate)
{
	if (ensure_init() || sc_check(sc, nqb, n_sites, tapered))
		return -1;
	const uint64_t total = (uint64_t)sc->Ma * sc->Mb;
	if (total == 0)
		return 0;
	hipLaunchKernelGGL(k_sc_scatter, dim3(grid_for(total, kBlock, 1u << 16)),
		dim3(kBlock), 0, g_stream, (double2 *)amp, sc->ma, sc->da,
		sc->mb, sc->db, (uint64_t)sc->Ma, (uint64_t)sc->Mb, n_sites,
		weight, tapered, accumulate);
	HIPCHK(hipGetLastError());
	return 0;
}

int ph2hip_sc_inner(const void *amp, uint32_t nqb, const struct ph2hip_sc *sc,
	uint32_t n_sites, double weight, int tapered, double *out)
{
	if (ensure_init() || sc_check(sc, nqb, n_sites, tapered) || ensure_red())
		return -1;
	const uint64_t total = (uint64_t)sc->Ma * sc->Mb;
	const uint32_t nb = grid_for(total, kBlock, kRedBlocks);
	double2 *partial = (double2 *)g_red;
	hipLaunchKernelGGL(k_sc_inner, dim3(nb), dim3(kBlock), 0, g_stream,
		(const double2 *)amp, sc->ma, sc->da, sc->mb, sc->db,
		(uint64_t)sc->Ma, (uint64_t)sc->Mb, n_sites, weight, tapered,
		partial);
	HIPCHK(hipGetLastError());
	hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(kBlock), 0, g_stream,
		partial, nb, partial + kRedBlocks);
	HIPCHK(hipGetLastError());
	return ph2hip_d2h(out, partial + kRedBlocks, sizeof(double2));
}

int ph2hip_prof_enable(int on)
{
	if (ensure_init())
		return -1;
	if (!on && g_pn && prof_drain())
		return -1;
	g_prof = on != 0;
	return 0;
}

int ph2hip_prof_read(struct ph2hip_prof *out, int reset)
{
	if (ensure_init() || prof_drain())
		return -1;
	*out = g_pacc;
	if (reset)
		g_pacc = ph2hip_prof{};
	return 0;
}

long ph2hip_prof_records(struct ph2hip_launch_rec *out, size_t cap,
	int reset)
{
	if (ensure_init() || prof_drain())
		return -1;
	size_t n = g_rec_n < cap ? g_rec_n : cap;
	const size_t skip = g_rec_n - n;   // keep the newest
	for (size_t i = 0; i < n; i++)
		out[i] = g_rec[(g_rec_head + skip + i) % PH2HIP_REC_MAX];
	if (reset)
		g_rec_n = g_rec_head = 0;
	return (long)n;
}

int ph2hip_set_fused_form(uint32_t pipe_mask, uint64_t persistent_wgs)
{
	if (ensure_init())
		return -1;
	pipe_init();
	g_pipe_mask = pipe_mask;
	if (persistent_wgs)
		g_persist_wgs = persistent_wgs;
	return 0;
}

int ph2hip_set_fused_direct(int on)
{
	g_direct = on != 0;
	return 0;
}

int ph2hip_set_fused_prefetch(uint32_t tile_bits, uint32_t amps,
	uint32_t wgs_per_cu)
{
	if (ensure_init())
		return -1;
	if (tile_bits != 12 && tile_bits != 13)
		return fail("set_fused_prefetch: tile %u has no half-tile form",
			tile_bits);
	if (amps != 0 && amps != 8 && amps != 16)
		return fail("set_fused_prefetch: %u amplitudes (0, 8 or 16)",
			amps);
	if (amps && (wgs_per_cu == 0 || wgs_per_cu > 16))
		return fail("set_fused_prefetch: %u workgroups per CU",
			wgs_per_cu);
	pipe_init();
	g_hpipe_pf[tile_bits] = amps;
	g_hpipe_wpc[tile_bits] = wgs_per_cu;
	return 0;
}

uint32_t ph2hip_fused_half_mask(void)
{
	return half_mask();
}

int ph2hip_set_fused_half(uint32_t tile_mask)
{
	half_mask();   // the environment default is read once, then replaced
	if (tile_mask & ~kHalfTiles)
		return fail("set_fused_half: mask 0x%x outside 0x%x", tile_mask,
			kHalfTiles);
	g_half_mask = tile_mask;
	return 0;
}

int ph2hip_mark(int slot)
{
	if (ensure_init())
		return -1;
	if (slot < 0 || slot >= 8)
		return fail("mark: slot %d", slot);
	if (!g_marks_made) {
		for (auto &e : g_marks)
			HIPCHK(hipEventCreate(&e));
		g_marks_made = true;
	}
	HIPCHK(hipEventRecord(g_marks[slot], g_stream));
	return 0;
}

int ph2hip_elapsed_ms(int a, int b, double *ms)
{
	if (!g_marks_made || a < 0 || a >= 8 || b < 0 || b >= 8)
		return fail("elapsed: no marks");
	float f = 0.f;
	HIPCHK(hipEventSynchronize(g_marks[b]));
	HIPCHK(hipEventElapsedTime(&f, g_marks[a], g_marks[b]));
	*ms = f;
	return 0;
}

int ph2hip_sample_cdf(size_t nstream, const double *const *cdf,
	const size_t *len, uint64_t *state, const size_t *ndraw, uint32_t *out)
{
	if (ensure_init())
		return -1;
	if (nstream == 0)
		return 0;
	if (nstream > (size_t)INT32_MAX)
		return fail("sample_cdf: %zu streams", nstream);
	size_t ntab = 0, nout = 0;
	for (size_t k = 0; k < nstream; k++) {
		if (len[k] == 0 || len[k] > UINT32_MAX || !cdf[k])
			return fail("sample_cdf: stream %zu: table of %zu "
				"entries", k, len[k]);
		if (!(state[4 * k] | state[4 * k + 1] | state[4 * k + 2] |
			    state[4 * k + 3]))
			return fail("sample_cdf: stream %zu: zero state", k);
		ntab += len[k];
		nout += ndraw[k];
	}
	DrawJob *hj = static_cast<DrawJob *>(malloc(sizeof(DrawJob) *
		nstream));
	if (!hj)
		return fail("sample_cdf: out of host memory");
	size_t to = 0, oo = 0;
	for (size_t k = 0; k < nstream; k++) {
		for (int i = 0; i < 4; i++)
			hj[k].s[i] = state[4 * k + i];
		hj[k].cdf_off = to;
		hj[k].len = len[k];
		hj[k].ndraw = ndraw[k];
		hj[k].out_off = oo;
		to += len[k];
		oo += ndraw[k];
	}
	double *dtab = nullptr;
	DrawJob *dj = nullptr;
	uint32_t *dout = nullptr;
	int rc = 0;
	if (hipMalloc(&dtab, sizeof(double) * ntab) != hipSuccess ||
	    hipMalloc(&dj, sizeof(DrawJob) * nstream) != hipSuccess ||
	    hipMalloc(&dout, sizeof(uint32_t) * (nout ? nout : 1)) !=
		    hipSuccess)
		rc = fail("sample_cdf: hipMalloc");
	for (size_t k = 0; k < nstream && !rc; k++)
		if (hipMemcpyAsync(dtab + hj[k].cdf_off, cdf[k],
			    sizeof(double) * len[k], hipMemcpyHostToDevice,
			    g_stream) != hipSuccess)
			rc = fail("sample_cdf: copy of table %zu", k);
	if (!rc && hipMemcpyAsync(dj, hj, sizeof(DrawJob) * nstream,
			   hipMemcpyHostToDevice, g_stream) != hipSuccess)
		rc = fail("sample_cdf: copy of the streams");
	if (!rc) {
		hipLaunchKernelGGL(k_cdf_draw, dim3((uint32_t)nstream), dim3(64),
			0, g_stream, dj, dtab, dout);
		if (hipGetLastError() != hipSuccess)
			rc = fail("sample_cdf: launch");
	}
	if (!rc && nout && hipMemcpyAsync(out, dout, sizeof(uint32_t) * nout,
			   hipMemcpyDeviceToHost, g_stream) != hipSuccess)
		rc = fail("sample_cdf: copy out");
	if (!rc && hipMemcpyAsync(hj, dj, sizeof(DrawJob) * nstream,
			   hipMemcpyDeviceToHost, g_stream) != hipSuccess)
		rc = fail("sample_cdf: copy of the states");
	if (hipStreamSynchronize(g_stream) != hipSuccess && !rc)
		rc = fail("sample_cdf: sync");
	if (!rc)
		for (size_t k = 0; k < nstream; k++)
			for (int i = 0; i < 4; i++)
				state[4 * k + i] = hj[k].s[i];
	(void)hipFree(dtab);
	(void)hipFree(dj);
	(void)hipFree(dout);
	free(hj);
	return rc;
}

int ph2hip_norm2(const void *amp, uint32_t nqb, double *out)
{
	if (ensure_init() || ensure_red())
		return -1;
	const uint64_t namp = UINT64_C(1) << nqb;
	const uint32_t nb = grid_for(namp, kBlock * 8, kRedBlocks);
	double2 *partial = (double2 *)g_red;
	hipLaunchKernelGGL(k_norm2, dim3(nb), dim3(kBlock), 0, g_stream,
		(const double2 *)amp, namp, partial);
	HIPCHK(hipGetLastError());
	hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(kBlock), 0, g_stream,
		partial, nb, partial + kRedBlocks);
	HIPCHK(hipGetLastError());
	double r[2];
	if (ph2hip_d2h(r, partial + kRedBlocks, sizeof r))
		return -1;
	*out = r[0];
	return 0;
}

int ph2hip_fill_hash(void *amp, uint64_t namp, uint64_t seed)
{
	if (ensure_init())
		return -1;
	hipLaunchKernelGGL(k_fill_hash, dim3(grid_for(namp, kBlock * 8,
		kMaxGrid)), dim3(kBlock), 0, g_stream, (double2 *)amp, namp,
		seed);
	HIPCHK(hipGetLastError());
	return 0;
}

int ph2hip_count_unequal(const void *a, const void *b, uint64_t namp,
	uint64_t *count)
{
	if (ensure_init() || ensure_red())
		return -1;
	unsigned long long *dc = (unsigned long long *)g_red;
	hipLaunchKernelGGL(k_fill, dim3(1), dim3(kBlock), 0, g_stream,
		(double2 *)g_red, (uint64_t)1);
	HIPCHK(hipGetLastError());
	hipLaunchKernelGGL(k_count_neq, dim3(grid_for(namp, kBlock * 8,
		kRedBlocks * 4)), dim3(kBlock), 0, g_stream,
		(const double2 *)a, (const double2 *)b, namp, dc);
	HIPCHK(hipGetLastError());
	unsigned long long c = 0;
	if (ph2hip_d2h(&c, dc, sizeof c))
		return -1;
	*count = c;
	return 0;
}


void ph2hip_batch_plan_free(struct ph2hip_batch_plan *pl)
{
	if (!pl)
		return;
	free(pl->desc);
	free(pl->items);
	free(pl->lstart);
	free(pl->ltiles);
	free(pl->lhalf);
	free(pl->lpass);
	free(pl);
}

int ph2hip_batch_plan_build(const struct ph2hip_batch_job *jobs, size_t njobs,
	uint32_t tile_bits, uint32_t low_bits, uint64_t namp_total,
	struct ph2hip_batch_plan **out)
{
	*out = nullptr;
	if (tile_bits < PH2HIP_TILE_MIN || tile_bits > PH2HIP_TILE_MAX)
		return fail("rotate_batched: tile_bits %u", tile_bits);
	size_t ndesc = 0, maxp = 0;
	for (size_t j = 0; j < njobs; j++) {
		const struct ph2hip_batch_job &J = jobs[j];
		if (J.nqb < tile_bits || J.nqb > 40 ||
		    J.amp_off + (UINT64_C(1) << J.nqb) > namp_total)
			return fail("rotate_batched: job %zu (nqb=%u) does not fit",
				j, J.nqb);
		size_t tot = 0;
		for (size_t p = 0; p < J.npass; p++) {
			if (J.pass_len[p] == 0 || J.pass_len[p] > PH2HIP_PASS_MAX)
				return fail("rotate_batched: job %zu pass %zu", j, p);
			tot += J.pass_len[p];
		}
		if (tot != J.n)
			return fail("rotate_batched: job %zu passes cover %zu of %zu",
				j, tot, J.n);
		for (size_t k = 0; k < J.n; k++)
			if (J.rot[k].x >> J.nqb)
				return fail("rotate_batched: job %zu term %zu", j, k);
		ndesc += J.npass;
		maxp = J.npass > maxp ? J.npass : maxp;
	}
	struct ph2hip_batch_plan *pl = (struct ph2hip_batch_plan *)calloc(1,
		sizeof *pl);
	if (!pl)
		return fail("rotate_batched: out of host memory");
	pl->ndesc = ndesc;
	pl->maxp = maxp;
	pl->tile_bits = tile_bits;
	pl->namp_total = namp_total;
	if (ndesc == 0) {
		*out = pl;
		return 0;
	}
	pl->desc = (FusedPass *)malloc(sizeof(FusedPass) * ndesc);
	pl->items = (BatchItem *)malloc(sizeof(BatchItem) * ndesc);
	pl->lstart = (size_t *)calloc(2 * maxp + 1, sizeof(size_t));
	pl->ltiles = (uint64_t *)calloc(2 * maxp, sizeof(uint64_t));
	pl->lhalf = (uint8_t *)calloc(2 * maxp, 1);
	pl->lpass = (uint32_t *)calloc(2 * maxp, sizeof(uint32_t));
	size_t *first = (size_t *)malloc(sizeof(size_t) * (njobs ? njobs : 1));
	int rc = 0;
	if (!pl->desc || !pl->items || !pl->lstart || !pl->ltiles ||
	    !pl->lhalf || !pl->lpass || !first) {
		rc = fail("rotate_batched: out of host memory");
	} else {
		// descriptors, job-major
		size_t d = 0;
		for (size_t j = 0; j < njobs && !rc; j++) {
			first[j] = d;
			size_t k0 = 0;
			for (size_t p = 0; p < jobs[j].npass && !rc; p++) {
				rc = 1;
				if (((half_mask() & kBatchHalfTiles) >> tile_bits) & 1)
					rc = build_pass_mode(pl->desc[d],
						jobs[j].rot + k0, jobs[j].pass_len[p],
						tile_bits, low_bits, k0, true);
				if (rc > 0)
					rc = build_pass_mode(pl->desc[d],
						jobs[j].rot + k0, jobs[j].pass_len[p],
						tile_bits, low_bits, k0, false);
				d++;
				k0 += jobs[j].pass_len[p];
			}
		}
	}
	if (!rc) {
		// work items, launch-major: pass p of every job that has one,
		// those in the half-tile form in one launch and the others in a
		// second (the two kernels differ in LDS size)
		size_t it = 0, l = 0;
		for (size_t p = 0; p < maxp; p++) {
			for (int h = 1; h >= 0; h--) {
				const size_t it0 = it;
				uint64_t t0 = 0;
				for (size_t j = 0; j < njobs; j++) {
					if (jobs[j].npass <= p)
						continue;
					const size_t dj = first[j] + p;
					if (((pl->desc[dj].dflags & 4u) != 0) != (h != 0))
						continue;
					pl->items[it].amp_off = jobs[j].amp_off;
					pl->items[it].tile0 = t0;
					pl->items[it].desc = dj;
					t0 += UINT64_C(1) << (jobs[j].nqb - tile_bits);
					it++;
				}
				if (it == it0)
					continue;
				pl->lstart[l] = it0;
				pl->ltiles[l] = t0;
				pl->lhalf[l] = (uint8_t)h;
				pl->lpass[l] = (uint32_t)p;
				l++;
			}
		}
		pl->lstart[l] = it;
		pl->nlaunch = l;
	}
	free(first);
	if (rc) {
		ph2hip_batch_plan_free(pl);
		return -1;
	}
	*out = pl;
	return 0;
}

int ph2hip_batch_plan_launch(void *amp, struct ph2hip_batch_plan *pl,
	int use_graph, uint64_t *launches)
{
	if (ensure_init())
		return -1;
	if (launches)
		*launches = 0;
	if (pl->ndesc == 0)
		return 0;
	const size_t ndesc = pl->ndesc, nlaunch = pl->nlaunch;
	const uint32_t tile_bits = pl->tile_bits;
	// The upload follows the work already queued on the stream, so the
	// launches of an earlier plan have read the buffers before they change.
	if (dev_reserve(&g_bdesc, &g_bdesc_cap, sizeof(FusedPass) * ndesc) ||
	    dev_reserve(&g_bitem, &g_bitem_cap, sizeof(BatchItem) * ndesc))
		return -1;
	if (hipMemcpyAsync(g_bdesc, pl->desc, sizeof(FusedPass) * ndesc,
		    hipMemcpyHostToDevice, g_stream) != hipSuccess ||
	    hipMemcpyAsync(g_bitem, pl->items, sizeof(BatchItem) * ndesc,
		    hipMemcpyHostToDevice, g_stream) != hipSuccess)
		return fail("rotate_batched: upload");
	// Profiling brackets every launch with events, so it launches
	// directly; otherwise the round graph runs the round.
	if (use_graph && !g_prof) {
		if (rg_launch(pl, (double2 *)amp))
			return -1;
		if (launches)
			*launches = nlaunch;
		return 0;
	}
	int rc = 0;
	for (size_t l = 0; l < nlaunch && !rc; l++) {
		const uint32_t ni = (uint32_t)(pl->lstart[l + 1] - pl->lstart[l]);
		const BatchItem *it = (const BatchItem *)g_bitem + pl->lstart[l];
		const FusedPass *dd = (const FusedPass *)g_bdesc;
		const bool hf = pl->lhalf[l] != 0;
		hipEvent_t stop = nullptr;
		ph2hip_launch_rec meta{};
		meta.tile_bits = tile_bits;
		meta.ngroups = ni;
		meta.direct = hf ? 4u : 0u;
		if (prof_begin(3, ni, pl->ltiles[l] << tile_bits, &stop,
			    &meta)) {
			rc = -1;
			break;
		}
		double2 *a = (double2 *)amp;
		const uint64_t nt = pl->ltiles[l];
		switch (tile_bits) {
		case 9: rc = launch_batch<9>(a, dd, it, ni, nt, hf); break;
		case 10: rc = launch_batch<10>(a, dd, it, ni, nt, hf); break;
		case 11: rc = launch_batch<11>(a, dd, it, ni, nt, hf); break;
		case 12: rc = launch_batch<12>(a, dd, it, ni, nt, hf); break;
		default: rc = launch_batch<13>(a, dd, it, ni, nt, hf); break;
		}
		if (!rc && prof_end(stop))
			rc = -1;
	}
	if (!rc && launches)
		*launches = nlaunch;
	return rc;
}

void ph2hip_round_graph_stats(uint64_t *builds, uint64_t *replays)
{
	*builds = g_rg_builds;
	*replays = g_rg_replays;
}

int ph2hip_rotate_batched(void *amp, uint64_t namp_total,
	const struct ph2hip_batch_job *jobs, size_t njobs, uint32_t tile_bits,
	uint32_t low_bits, int use_graph, uint64_t *launches)
{
	if (ensure_init())
		return -1;
	struct ph2hip_batch_plan *pl = nullptr;
	if (ph2hip_batch_plan_build(jobs, njobs, tile_bits, low_bits, namp_total,
		    &pl) < 0)
		return -1;
	int rc = ph2hip_batch_plan_launch(amp, pl, use_graph, launches);
	if (hipStreamSynchronize(g_stream) != hipSuccess && !rc)
		rc = fail("rotate_batched: sync");
	ph2hip_batch_plan_free(pl);
	return rc;
}

int ph2hip_scatter_set(void *amp, uint64_t namp_total, const uint64_t *idx,
	const double *val, size_t n)
{
	if (ensure_init())
		return -1;
	if (n == 0)
		return 0;
	for (size_t k = 0; k < n; k++)
		if (idx[k] >= namp_total)
			return fail("scatter_set: index %llu beyond %llu",
				(unsigned long long)idx[k],
				(unsigned long long)namp_total);
	// later entries win: keep the last occurrence of every index
	size_t *ord = (size_t *)malloc(sizeof(size_t) * n);
	uint64_t *ui = (uint64_t *)malloc(sizeof(uint64_t) * n);
	double *uv = (double *)malloc(sizeof(double) * 2 * n);
	if (!ord || !ui || !uv) {
		free(ord);
		free(ui);
		free(uv);
		return fail("scatter_set: out of host memory");
	}
	for (size_t k = 0; k < n; k++)
		ord[k] = k;
	std::stable_sort(ord, ord + n,
		[&](size_t x, size_t y) { return idx[x] < idx[y]; });
	size_t m = 0;
	for (size_t k = 0; k < n; k++) {
		if (k + 1 < n && idx[ord[k + 1]] == idx[ord[k]])
			continue;
		ui[m] = idx[ord[k]];
		uv[2 * m] = val[2 * ord[k]];
		uv[2 * m + 1] = val[2 * ord[k] + 1];
		m++;
	}
	uint64_t *di = nullptr;
	double2 *dv = nullptr;
	int rc = 0;
	if (hipMalloc(&di, sizeof(uint64_t) * m) != hipSuccess ||
	    hipMalloc(&dv, sizeof(double2) * m) != hipSuccess)
		rc = fail("scatter_set: hipMalloc");
	if (!rc && (hipMemcpyAsync(di, ui, sizeof(uint64_t) * m,
			    hipMemcpyHostToDevice, g_stream) != hipSuccess ||
		    hipMemcpyAsync(dv, uv, sizeof(double2) * m,
			    hipMemcpyHostToDevice, g_stream) != hipSuccess))
		rc = fail("scatter_set: copy");
	if (!rc) {
		hipLaunchKernelGGL(k_scatter_set, dim3(grid_for(m, kBlock, kMaxGrid)),
			dim3(kBlock), 0, g_stream, (double2 *)amp, di, dv,
			(uint64_t)m);
		if (hipGetLastError() != hipSuccess)
			rc = fail("scatter_set: launch");
	}
	if (hipStreamSynchronize(g_stream) != hipSuccess && !rc)
		rc = fail("scatter_set: sync");
	(void)hipFree(di);
	(void)hipFree(dv);
	free(ord);
	free(ui);
	free(uv);
	return rc;
}

int ph2hip_gather_abs(const void *amp, uint64_t namp_total,
	const uint64_t *idx, size_t n, double *out)
{
	for (size_t k = 0; k < n; k++)
		if (idx[k] >= namp_total)
			return fail("gather: index beyond the buffer");
	// ph2hip_gather checks against 2^nqb; pass the buffer's bit width
	uint32_t bits = 0;
	while ((UINT64_C(1) << bits) < namp_total)
		bits++;
	return ph2hip_gather(amp, bits, idx, n, out);
}

long ph2hip_pass_groups(const struct ph2hip_rot *rot, size_t n,
	uint32_t tile_bits, uint32_t low_bits, uint32_t *out, size_t cap)
{
	if (n == 0 || n > PH2HIP_PASS_MAX || tile_bits < PH2HIP_TILE_MIN ||
	    tile_bits > PH2HIP_TILE_MAX)
		return -1;
	FusedPass P;
	if (build_pass(P, rot, n, tile_bits, low_bits, 0))
		return -1;
	if (P.ngroups > cap)
		return fail("pass_groups: %u groups, room for %zu", P.ngroups,
			cap);
	for (uint32_t gi = 0; gi < P.ngroups; gi++) {
		const FusedGroup &G = P.grp[gi];
		uint32_t *o = out + (size_t)gi * PH2HIP_GROUP_WORDS;
		for (int i = 0; i < 4; i++)
			o[i] = G.u[i];
		for (int i = 0; i < 12; i++)
			o[4 + i] = G.jpos[i];
		o[16] = G.xh;
		o[17] = G.xk;
		o[18] = G.first;
		o[19] = G.count;
	}
	return P.ngroups;
}

int ph2hip_pass_shape(const struct ph2hip_rot *rot, size_t n,
	uint32_t tile_bits, uint32_t low_bits, uint32_t *shape)
{
	if (n == 0 || n > PH2HIP_PASS_MAX || tile_bits < PH2HIP_TILE_MIN ||
	    tile_bits > PH2HIP_TILE_MAX)
		return -1;
	FusedPass P;
	if (build_pass(P, rot, n, tile_bits, low_bits, 0))
		return -1;
	shape[0] = P.L;
	shape[1] = P.g;
	shape[2] = P.ngroups;
	shape[3] = P.dflags;
	return 0;
}

} // extern "C"

