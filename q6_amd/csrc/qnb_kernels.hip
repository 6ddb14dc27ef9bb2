// qnb_kernels.hip -- HIP kernels of the nonbonded engine for gfx950.
//
// Work decomposition is receiver-centric: one wavefront owns one charge group
// (a solute group or a water) and computes every interaction of its atoms with
// all partners, so pair forces need no atomics and the result does not depend
// on scheduling.  Each unordered pair is evaluated from both sides; its energy
// is counted once, on the side that owns the pair under the reference's parity
// rule (nonbondene.f90:1589-1591), which also decides which side's list the
// pair is exported in.  Q-atoms, whose partners are all waters and solute
// atoms within Rcq, are the one many-to-one case: their forces are summed per
// wavefront and added with one atomic per wave and component.
#include "qnb_dev.h"

namespace qnb {

__device__ __forceinline__ int group_ref_atom(const DevSys& S, int g) {
  return g < S.ncgp_solute ? ldg(S.gswitch + g) : S.nat_solute + S.sa * (g - S.ncgp_solute);
}

// Group test of the iuse_switch_atom = 0 builders for a pair with a solute
// group: true when any atom pair (solute-solute, nbpplis2(_lrf)
// nonbondene.f90:950-1089, 1369-1488) or any solute atom and the water's
// first atom (nbpwlis2(_lrf) :2137-2230, 2493-2620) lies within cut2;
// otherwise r2 is the squared distance of the last pair the reference scans
// (the groups' last atoms; symmetric, so either group may be the receiver),
// which its _lrf forms compare with RcLRF^2
__device__ __forceinline__ bool any_atom_pair(const DevSys& S, const double* __restrict__ x, int a, int b, bool a_sol,
                                            bool b_sol, int ra, int rb, double cut2, double& r2) {
  r2 = 1.0;
  if (a_sol && b_sol) {
    const int a1 = ldg(S.glast + a), b0 = ldg(S.gfirst + b), b1 = ldg(S.glast + b);
    for (int ia = ldg(S.gfirst + a); ia <= a1; ia++) {
      const int i = ldg(S.cgpatom + ia);
      for (int jb = b0; jb <= b1; jb++) {
        r2 = dist4(x, i, ldg(S.cgpatom + jb));
        if (r2 <= cut2) return true;
      }
    }
    return false;
  }
  const int g = a_sol ? a : b, w = a_sol ? rb : ra;
  const int g1 = ldg(S.glast + g);
  for (int k = ldg(S.gfirst + g); k <= g1; k++) {
    r2 = dist4(x, ldg(S.cgpatom + k), w);
    if (r2 <= cut2) return true;
  }
  return false;
}

// First water (index order) within the cut-off of each solute group's switch
// atom, for the LRF-off switch-atom solute-water list: nbpwlist
// (nonbondene.f90:2665-2691) reuses its switch-atom variable `ia` as the loop
// index over the group's atoms, so once a water is inside, the distance test
// of every later water is made from atom cgp(ig)%last + 1.  One wavefront per
// solute group; excluded groups and waters are skipped as the reference does.
__global__ __launch_bounds__(256) void k_pw_first(const DevSys* __restrict__ systems) {
  const DevSys& S = systems[blockIdx.y];
  const int g = blockIdx.x * (blockDim.x / QNB_WAVE) + threadIdx.x / QNB_WAVE;
  if (g >= S.ncgp_solute || !pw_quirk(S)) return;
  const int lane = threadIdx.x % QNB_WAVE;
  const int ia = ldg(S.gswitch + g);
  int first = S.nwat;
  if (!ldg(S.excl + ia)) {
    for (int w0 = 0; w0 < S.nwat; w0 += QNB_WAVE) {
      const int w = w0 + lane;
      bool in = false;
      if (w < S.nwat) {
        const int ja = S.nat_solute + S.sa * w;
        in = !ldg(S.excl + ja) && dist4(S.x, ia, ja) <= S.rcpw2;
      }
      const uint64_t m = __ballot(in);
      if (m) {
        first = w0 + __ffsll((long long)m) - 1;
        break;
      }
    }
  }
  if (lane == 0) S.pw_first[g] = first;
}

// the owner side of a charge-group pair (a != b) under the parity rule
__device__ __forceinline__ bool owns(int a, int b) {
  return (a < b && ((a + b) & 1) == 0) || (a > b && ((a + b) & 1) == 1);
}

// ---------------------------------------------------------------------------
// LRF expansion (lrf_update, nonbondene.f90:628-725), restated on the 20
// independent components of the symmetric tensors: acc[0] phi0, acc[1..3]
// phi1, acc[4..9] phi2 (xx xy xz yy yz zz), acc[10..19] phi3 (xxx xxy xxz xyy
// xyz xzz yyy yyz yzz zzz).  dr = x_i - centre in FP64, terms in R; the
// caller sums one source group's terms in R (at most a group's atoms) and
// adds the group's partial to its FP64 accumulators.
template <typename R>
__device__ __forceinline__ void lrf_accumulate(R* acc, double drx, double dry, double drz, double q) {
  R dx = (R)drx, dy = (R)dry, dz = (R)drz;
  R r2 = dx * dx + dy * dy + dz * dz;
  R f0, f1, f2;
  if constexpr (sizeof(R) == 4) {
    R ri2, ri;
    rinv(r2, ri2, ri);
    f0 = (R)q * ri * ri2;
    f1 = R(3) * f0 * ri2;
    f2 = -f1 * ri2;
  } else {
    f0 = (R)q / (r2 * sqrt(r2));
    f1 = R(3) * f0 / r2;
    f2 = -f1 / r2;
  }
  acc[0] += f0 * r2;
  acc[1] -= dx * f0;
  acc[2] -= dy * f0;
  acc[3] -= dz * f0;
  R xx = dx * dx, xy = dx * dy, xz = dx * dz, yy = dy * dy, yz = dy * dz, zz = dz * dz;
  acc[4] += xx * f1 - f0;
  acc[5] += xy * f1;
  acc[6] += xz * f1;
  acc[7] += yy * f1 - f0;
  acc[8] += yz * f1;
  acc[9] += zz * f1 - f0;
  R g = R(5) * f2, h = f2 * r2;
  R gx = g * dx, gy = g * dy, gz = g * dz;
  acc[10] += gx * xx - R(3) * h * dx;
  acc[11] += gx * xy - h * dy;
  acc[12] += gx * xz - h * dz;
  acc[13] += gx * yy - h * dx;
  acc[14] += gx * yz;
  acc[15] += gx * zz - h * dx;
  acc[16] += gy * yy - R(3) * h * dy;
  acc[17] += gy * yz - h * dz;
  acc[18] += gy * zz - h * dy;
  acc[19] += gz * zz - R(3) * h * dz;
}

// FP32 form for the mixed path: dr already formed in FP32 from the packed
// positions (x_i and the centre each rounded once), q the atom's charge.
// The lane keeps raw sums that need no per-atom constant or subtraction:
//   p[0] = sum q/r,  p[1..3] = sum f0 d_a,  p[4..9] = sum w1 d_a d_b,
//   p[10..19] = sum w2 d_a d_b d_c
// with f0 = q/r^3, w1 = f0/r^2, w2 = w1/r^2 (component order as acc).
// lrf_update's field1 = 3 f0/r^2 and its 5 field2 d d d term (field2 =
// -field1/r^2) give phi2_ab = 3 p_ab - delta_ab sum f0 and
// phi3_abc = -15 p_abc + 3 (delta_ab H_c + delta_ac H_b + delta_bc H_a),
// H_a = sum w1 d_a; as w1 r^2 = f0 and w2 r^2 = w1, sum f0 is the trace of
// p[4..9] and H_a the trace of p[10..19] over the other two indices, which
// lrf_flush_f forms in FP64.
#define LRF_NF 20
typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d4v __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void lrf_accumulate_f(float* p, float dx, float dy, float dz, float q) {
  const float r2 = dx * dx + dy * dy + dz * dz;
  float ri2, ri;
  rinv(r2, ri2, ri);
  const float qri = q * ri;
  p[0] += qri;
  const float f0 = qri * ri2;
  const float w1 = f0 * ri2;
  const float w2 = w1 * ri2;
  p[1] += dx * f0;
  p[2] += dy * f0;
  p[3] += dz * f0;
  const float xx = dx * dx, xy = dx * dy, xz = dx * dz, yy = dy * dy, yz = dy * dz, zz = dz * dz;
  p[4] += xx * w1;
  p[5] += xy * w1;
  p[6] += xz * w1;
  p[7] += yy * w1;
  p[8] += yz * w1;
  p[9] += zz * w1;
  const float gx = w2 * dx, gy = w2 * dy, gz = w2 * dz;
  p[10] += gx * xx;
  p[11] += gx * xy;
  p[12] += gx * xz;
  p[13] += gx * yy;
  p[14] += gx * yz;
  p[15] += gx * zz;
  p[16] += gy * yy;
  p[17] += gy * yz;
  p[18] += gy * zz;
  p[19] += gz * zz;
}

// adds a lane's raw FP32 sums to its FP64 expansion
__device__ __forceinline__ void lrf_flush_f(double* acc, const float* p) {
  double q[LRF_NF];
#pragma unroll
  for (int k = 0; k < LRF_NF; k++) q[k] = (double)p[k];
  const double s0 = q[4] + q[7] + q[9];
  const double hx = q[10] + q[13] + q[15], hy = q[11] + q[16] + q[18], hz = q[12] + q[17] + q[19];
  acc[0] += q[0];
  acc[1] -= q[1];
  acc[2] -= q[2];
  acc[3] -= q[3];
  acc[4] += 3.0 * q[4] - s0;
  acc[5] += 3.0 * q[5];
  acc[6] += 3.0 * q[6];
  acc[7] += 3.0 * q[7] - s0;
  acc[8] += 3.0 * q[8];
  acc[9] += 3.0 * q[9] - s0;
  acc[10] += 9.0 * hx - 15.0 * q[10];
  acc[11] += 3.0 * hy - 15.0 * q[11];
  acc[12] += 3.0 * hz - 15.0 * q[12];
  acc[13] += 3.0 * hx - 15.0 * q[13];
  acc[14] -= 15.0 * q[14];
  acc[15] += 3.0 * hx - 15.0 * q[15];
  acc[16] += 9.0 * hy - 15.0 * q[16];
  acc[17] += 3.0 * hz - 15.0 * q[17];
  acc[18] += 3.0 * hy - 15.0 * q[18];
  acc[19] += 9.0 * hz - 15.0 * q[19];
}

// Raw buffer loads (32-bit byte offsets from a base in scalar registers; an
// offset at or past `bytes` reads 0): the sweep's packed arrays are indexed
// without 64-bit address arithmetic or clamping.
typedef unsigned int u4v __attribute__((vector_size(16)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, size_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ f4v buf_f4(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return __builtin_bit_cast(f4v, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}
__device__ __forceinline__ d4v buf_d4(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  const d2v lo = __builtin_bit_cast(d2v, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
  const d2v hi = __builtin_bit_cast(d2v, __builtin_amdgcn_raw_buffer_load_b128(r, off + 16, 0, 0));
  return d4v{lo.x, lo.y, hi.x, hi.y};
}
__device__ __forceinline__ int buf_i(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  return (int)__builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
}

// Packs the LRF sources before a list update, one record per atom in
// cgpatom order (a charge group's atoms are one run): x, y, z and the charge,
// 0 for Q-atoms (lrf_update skips them); FP32 for the mixed path (each
// coordinate rounded once), FP64 for the plain one.  Also packs each group's
// reference atom and its exclusion flag for the group tests of the sweep.
__global__ __launch_bounds__(256) void k_pack_xq(const DevSys* __restrict__ systems) {
  const DevSys& S = systems[blockIdx.y];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= S.ncgpat) return;
  const int i = ldg(S.cgpatom + k);
  const auto x = gp(S.x);
  const double q = ldg(S.iqatom + i) != 0 ? 0.0 : ldg(S.crg + i);
  if (S.xqg) {
    const auto o = gpm(S.xqg);
    o[4 * k] = (float)x[3 * i];
    o[4 * k + 1] = (float)x[3 * i + 1];
    o[4 * k + 2] = (float)x[3 * i + 2];
    o[4 * k + 3] = (float)q;
  }
  if (S.xqgd) {
    const auto o = gpm(S.xqgd);
    o[4 * k] = x[3 * i];
    o[4 * k + 1] = x[3 * i + 1];
    o[4 * k + 2] = x[3 * i + 2];
    o[4 * k + 3] = q;
  }
  if (k < S.ncgp) {  // thread k also packs group k's reference atom
    const int rb = group_ref_atom(S, k);
    const auto o = gpm(S.gref);
    o[4 * k] = x[3 * rb];
    o[4 * k + 1] = x[3 * rb + 1];
    o[4 * k + 2] = x[3 * rb + 2];
    o[4 * k + 3] = ldg(S.excl + rb) ? 1.0 : 0.0;
  }
}

// index of the symmetric third-rank component (a<=b<=c sorted) in acc[10..]
__device__ __forceinline__ int t3index(int a, int b, int c) {
  // sort three indices
  if (a > b) { int t = a; a = b; b = t; }
  if (b > c) { int t = b; b = c; c = t; }
  if (a > b) { int t = a; a = b; b = t; }
  // lexicographic rank among (000 001 002 011 012 022 111 112 122 222)
  const int base[3] = {0, 6, 9};
  if (a == 0) {
    const int r0[3][3] = {{0, 1, 2}, {0, 3, 4}, {0, 0, 5}};
    return 10 + r0[b][c];
  }
  if (a == 1) return 10 + (b == 1 ? (c == 1 ? 6 : 7) : 8);
  (void)base;
  return 19;
}

// one packed LRF source record: FP32 on the mixed path, FP64 on the plain one
template <typename R>
__device__ __forceinline__ auto lrf_rec(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  if constexpr (sizeof(R) == 4) return buf_f4(r, off);
  else return buf_d4(r, off);
}

// dist4 on packed reference positions: |b - a|^2, ((dx^2 + dy^2) + dz^2)
__device__ __forceinline__ double dist4v(const d4v& a, const d4v& b) {
#pragma clang fp contract(off)
  const double dx = b.x - a.x, dy = b.y - a.y, dz = b.z - a.z;
  double r = dx * dx;
  r = r + dy * dy;
  r = r + dz * dz;
  return r;
}

// ---------------------------------------------------------------------------
// Pair-list update with LRF (make_pair_lists, nonbondene.f90:749-837, SPH):
// one wavefront per receiver charge group a, sweeping the other groups in
// chunks of 64 (lane = group b; reference atoms are L2-resident).  A pair
// inside its class cut-off goes to a's neighbour row through a ballot/prefix
// compaction; the pairs beyond it (and inside RcLRF) form the chunk's far
// word, and the wave then adds the charges of those groups' atoms to a's
// expansion lane by atom over the chunk's run of packed records (lrf_update,
// nonbondene.f90:628-725), so every lane has an atom whatever the group sizes.
// ANY: some system of the launch has iuse_switch_atom = 0 (any_atom), whose
// group tests scan atom pairs; the switch-atom form is compiled without it.
template <typename R, bool ANY>
__global__ __launch_bounds__(512) void k_lists_lrf(const DevSys* __restrict__ systems) {
  const DevSys& S = systems[blockIdx.y];
  const int a = blockIdx.x * (blockDim.x / QNB_WAVE) + threadIdx.x / QNB_WAVE;
  if (a >= S.ncgp) return;
  const int lane = threadIdx.x % QNB_WAVE;
  const double* __restrict__ x = S.x;
  const int ra = group_ref_atom(S, a);
  const bool a_sol = a < S.ncgp_solute;
  // centre of a (cgp_centers, nonbondene.f90:43-78): sum in atom order, then divide
  double cx = 0.0, cy = 0.0, cz = 0.0;
  const int ga0 = ldg(S.gfirst + a), ga1 = ldg(S.glast + a);
  for (int k = ga0; k <= ga1; k++) {
    int at = ldg(S.cgpatom + k);
    cx = cx + ldg(x + 3 * at);
    cy = cy + ldg(x + 3 * at + 1);
    cz = cz + ldg(x + 3 * at + 2);
  }
  double nat = (double)(ga1 - ga0 + 1);
  cx = cx / nat;
  cy = cy / nat;
  cz = cz / nat;
  const float cxf = (float)cx, cyf = (float)cy, czf = (float)cz;
  double acc[20];
#pragma unroll
  for (int k = 0; k < 20; k++) acc[k] = 0.0;
  int count = 0;
  int* __restrict__ row = S.nbr + (size_t)a * S.cap;
  // mixed path: the lane's raw FP32 sums since the last flush to the FP64
  // expansion, flushed after a chunk once a lane has been offered LRF_FLUSH
  // atoms (at most LRF_FLUSH - 1 plus one chunk's share)
  constexpr int LRF_FLUSH = 16;
  float pf[LRF_NF];
#pragma unroll
  for (int k = 0; k < LRF_NF; k++) pf[k] = 0.0f;
  int offered = 0;
  if (!ldg(S.excl + ra)) {
    const auto grs = buf_rsrc(S.gref, (size_t)S.ncgp * 32);
    const d4v A = buf_d4(grs, (uint32_t)a * 32);
    const auto rs = buf_rsrc(sizeof(R) == 4 ? (const void*)S.xqg : (const void*)S.xqgd,
                             (size_t)S.ncgpat * (sizeof(R) == 4 ? 16 : 32));
    const auto gs = buf_rsrc(S.gslot, (size_t)S.ncgpat * 4);
    // the cut-off of a pair with a solute group and with a water, given a
    const double cut_sol = a_sol ? S.rcpp2 : S.rcpw2, cut_wat = a_sol ? S.rcpw2 : S.rcww2;
    const bool quirk = pw_quirk(S), any_atom = ANY && S.any_atom, use_lrf = S.use_lrf;
    const double rclrf2 = S.rclrf2;
    for (int b0 = 0; b0 < S.ncgp; b0 += QNB_WAVE) {
      const int b = b0 + lane;
      const d4v G = buf_d4(grs, (uint32_t)b * 32);  // 0 past the last group
      bool in = false, far = false;
      if (b < S.ncgp && b != a && G.w == 0.0) {
        const bool b_sol = b < S.ncgp_solute;
        const double cut2 = b_sol ? cut_sol : cut_wat;
        double r2;
        bool inside;
        if (any_atom && (a_sol || b_sol)) {
          inside = any_atom_pair(S, x, a, b, a_sol, b_sol, ra, group_ref_atom(S, b), cut2, r2);
        } else if (a_sol != b_sol && quirk) {
          // nbpwlist's test point: the switch atom up to the group's first
          // water inside, atom glast + 1 after it (LRF off: r2 unused)
          const int g = a_sol ? a : b, w = (a_sol ? b : a) - S.ncgp_solute;
          const int wf = ldg(S.pw_first + g);
          r2 = 0.0;
          inside = w == wf ||
                   (w > wf && dist4(x, ldg(S.glast + g) + 1, a_sol ? group_ref_atom(S, b) : ra) <= cut2);
        } else {
          r2 = dist4v(A, G);
          inside = r2 <= cut2;
        }
        in = inside;
        far = !inside && use_lrf && r2 <= rclrf2;
      }
      const uint64_t m = __ballot(in);
      const int pos = count + __popcll(m & ((1ull << lane) - 1ull));
      if (in && pos < S.cap) stg(row + pos, b);
      count += __popcll(m);
      const uint64_t fw = __ballot(far);
      if (fw != 0) {
        // the chunk's atoms: the run from b0's first to its last group's last
        const int k0 = ldg(S.gfirst + b0), k1 = ldg(S.glast + min(b0 + QNB_WAVE, S.ncgp) - 1) + 1;
        // one iteration ahead: the next record and its group slot are in
        // flight while this one is evaluated (reads past the run are unused)
        constexpr uint32_t RB = sizeof(R) == 4 ? 16 : 32;
        int k = k0 + lane;
        uint32_t off = (uint32_t)k * RB, offs = (uint32_t)k * 4;
        auto r = lrf_rec<R>(rs, off);
        int sl = buf_i(gs, offs);
        for (; k - lane < k1; k += QNB_WAVE) {
          off += QNB_WAVE * RB;
          offs += QNB_WAVE * 4;
          const auto rn = lrf_rec<R>(rs, off);
          const int sn = buf_i(gs, offs);
          if (k < k1 && ((fw >> sl) & 1ull)) {
            if constexpr (sizeof(R) == 4) lrf_accumulate_f(pf, r.x - cxf, r.y - cyf, r.z - czf, r.w);
            else lrf_accumulate<double>(acc, r.x - cx, r.y - cy, r.z - cz, r.w);
          }
          r = rn;
          sl = sn;
        }
        offered += (k1 - k0 + QNB_WAVE - 1) / QNB_WAVE;
      }
      if constexpr (sizeof(R) == 4) {
        if (offered >= LRF_FLUSH) {
          lrf_flush_f(acc, pf);
#pragma unroll
          for (int c = 0; c < LRF_NF; c++) pf[c] = 0.0f;
          offered = 0;
        }
      }
    }
  }
  if constexpr (sizeof(R) == 4) lrf_flush_f(acc, pf);
#pragma unroll
  for (int k = 0; k < 20; k++) acc[k] = wave_sum_all(acc[k]);
  if (lane == 0) {
    S.nnbr[a] = count;
    if (count > S.cap) atomicMax(S.need, count);
    double* L = S.lrf + (size_t)QNB_NLRF * a;
    L[0] = cx;
    L[1] = cy;
    L[2] = cz;
    L[3] = acc[0];
    L[4] = acc[1];
    L[5] = acc[2];
    L[6] = acc[3];
    const int p2[3][3] = {{4, 5, 6}, {5, 7, 8}, {6, 8, 9}};
    for (int k = 0; k < 3; k++)
      for (int c = 0; c < 3; c++) L[7 + 3 * k + c] = acc[p2[k][c]];
    for (int n = 0; n < 3; n++)
      for (int j = 0; j < 3; j++)
        for (int c = 0; c < 3; c++) L[16 + 3 * (j + 3 * n) + c] = acc[t3index(n, j, c)];
  }
}

// ---------------------------------------------------------------------------
// Q-atom lists (nbqplist nonbondene.f90:3647-3742, nbqwlist 3862-3944): flags
// and ordered compaction of the waters and solute atoms within Rcq of xpcent.
// One block of 1024 threads per system; each thread takes a contiguous chunk
// so the compacted order is the reference's append order.
__global__ __launch_bounds__(1024) void k_qlists(const DevSys* __restrict__ systems) {
  const DevSys& S = systems[blockIdx.x];
  __shared__ int scan[1024];
  const int t = threadIdx.x, nt = blockDim.x;
  // waters
  {
    const int per = (S.nwat + nt - 1) / nt, w0 = t * per, w1 = min(S.nwat, w0 + per);
    int c = 0;
    for (int w = w0; w < w1; w++) {
      int o = S.nat_solute + S.sa * w;
      bool in = !S.excl[o] && dist4p(S.x, o, S.xp) <= S.rcq2;
      S.qw_in[w] = in;
      c += in;
    }
    scan[t] = c;
    __syncthreads();
    for (int off = 1; off < nt; off <<= 1) {
      int v = t >= off ? scan[t - off] : 0;
      __syncthreads();
      scan[t] += v;
      __syncthreads();
    }
    int pos = scan[t] - c;
    for (int w = w0; w < w1; w++)
      if (S.qw_in[w]) S.qw_list[pos++] = w;
    if (t == nt - 1) S.nq_lists[0] = scan[t];
    __syncthreads();
  }
  // solute atoms, group by group
  {
    const int per = (S.ncgp_solute + nt - 1) / nt, g0 = t * per, g1 = min(S.ncgp_solute, g0 + per);
    int c = 0;
    for (int g = g0; g < g1; g++) {
      const bool in = qp_group_in(S, S.x, g);
      S.qp_in[g] = in;
      if (in)
        for (int k = S.gfirst[g]; k <= S.glast[g]; k++) c += !S.qconn3[S.cgpatom[k]];
    }
    scan[t] = c;
    __syncthreads();
    for (int off = 1; off < nt; off <<= 1) {
      int v = t >= off ? scan[t - off] : 0;
      __syncthreads();
      scan[t] += v;
      __syncthreads();
    }
    int pos = scan[t] - c;
    for (int g = g0; g < g1; g++)
      if (S.qp_in[g])
        for (int k = S.gfirst[g]; k <= S.glast[g]; k++) {
          int i = S.cgpatom[k];
          if (!S.qconn3[i]) S.qp_list[pos++] = i;
        }
    if (t == nt - 1) S.nq_lists[1] = scan[t];
  }
}

// term bits (include/qnb.h enum qnb_term)
enum : uint32_t { T_WW = QNB_T_WW, T_QW = QNB_T_QW, T_PW = QNB_T_PW, T_PP = QNB_T_PP, T_QP = QNB_T_QP,
                  T_LRF = QNB_T_LRF, T_QQ = QNB_T_QQ, T_QQP = QNB_T_QQP, T_RAD = QNB_T_RAD, T_POL = QNB_T_POL,
                  T_SINT = QNB_T_SINT };

// ---------------------------------------------------------------------------
// Receiver-side forces: one wavefront per charge group a.  The wave expands
// its neighbour row into partner atoms 64 groups at a time (a wave prefix sum
// over the groups' atom counts, staged in LDS) and gives every lane one
// partner atom at a time, so lanes stay balanced however the row mixes 3-atom
// waters and larger solute groups.  A lane forms the partner's position
// relative to the receiver's reference atom once in FP64, rounds it to R and
// evaluates it against the receiver atoms (offsets from the same reference
// point), accumulating per-lane forces in R; the wave reduces them in FP64.
// Pair forms: water-water as nonbond_ww_spc / nonbond_ww
// (nonbondene.f90:5886-5975, 5509-5573: O-O LJ + Coulomb, the other 8
// Coulomb only for SPC), pw from the pw table (nonbond_pw 4856-4916), pp from
// pp_param (nonbond_pp 4694-4760).  Writes (not adds) d for every atom of a,
// and the wave's energy partials into epart[a] on the sampled steps.

// Partner staging: the wave takes up to 64 neighbour groups whose atoms fit
// in FC_CAP slots (wave prefix sum over the groups' atom counts; groups
// larger than FC_CAP are refused at set-up) and writes, per partner atom,
// its index, its position relative to the receiver's reference atom (formed
// in FP64, rounded to R) and a tag: bits 0-1 the atom's rank in its group,
// bit 2 solute partner, bit 3 the receiver owns the group pair's energy,
// bit 4 a Q-atom (its pairs belong to the Q kernels).  Returns the number of
// staged atoms; *ngroups receives the number of groups consumed.
#define FT_SOL 4
#define FT_OWN 8
#define FT_Q 16
#define FT_FAR 32  // solute partner group that holds no excluded or 1-4 pair with the receiver group
// pp_param from data at hand: the two atoms' types, charges and exclusion
// masks (bit 31: partners beyond the 25-atom window, looked up in the long
// list) and the type tables (LJ factors, ljcod) in LDS.  Same result as
// pp_param.
// the pair (lo, hi) in lo's long exclusion / 1-4 list: 1 excluded, 2 1-4,
// 0 neither (kept out of line: rare, and its loop would stop the callers'
// receiver-atom loops from unrolling)
__device__ __noinline__ int long_pair_kind(const DevSys& S, int lo, int hi) {
  bool ex = false, is14 = false;
  for (int m = ldg(S.lx_off + lo); m < ldg(S.lx_off + lo + 1); m++)
    if (ldg(S.lx_j + m) == hi) {
      if (ldg(S.lx_kind + m) == 0) ex = true;
      else is14 = true;
    }
  return ex ? 1 : (is14 ? 2 : 0);
}

struct PPAtom {
  int i, t;
  uint32_t xm, m14;
  double q;
};
__device__ __forceinline__ bool pp_param_fast(const DevSys& S, const PPAtom& P, const PPAtom& Q,
                                              const double* __restrict__ avdw, const double* __restrict__ bvdw,
                                              const uint8_t* __restrict__ ljc, double& A, double& B, double& el) {
  const bool pl = P.i < Q.i;
  const int dk = pl ? Q.i - P.i : P.i - Q.i;
  const uint32_t xm = pl ? P.xm : Q.xm, m14 = pl ? P.m14 : Q.m14;
  bool is14 = false;
  if (dk <= 25) {
    const uint32_t bit = 1u << (dk - 1);
    if (xm & bit) return false;
    is14 = (m14 & bit) != 0;
  } else if (xm >> 31) {
    const int lo = pl ? P.i : Q.i, hi = pl ? Q.i : P.i;
    const int kind = long_pair_kind(S, lo, hi);
    if (kind == 1) return false;
    is14 = kind == 2;
  }
  const int vdw = is14 ? 3 : ljc[P.t + Q.t * S.ntypes];
  const int k = vdw - 1;
  if (S.ivdw_rule == 1) {
    A = avdw[P.t * 3 + k] * avdw[Q.t * 3 + k];
    B = bvdw[P.t * 3 + k] * bvdw[Q.t * 3 + k];
  } else {
    double tA = avdw[P.t * 3 + k] + avdw[Q.t * 3 + k];
    tA = tA * tA;
    tA = tA * tA * tA;
    double tB = bvdw[P.t * 3 + k] * bvdw[Q.t * 3 + k];
    A = (tA * tA) * tB;
    B = 2.0 * tA * tB;
  }
  el = P.q * Q.q;
  const bool set = fabs(A) > 1.0e-10 || fabs(B) > 1.0e-10 || fabs(el) > 1.0e-10;
  if (vdw == 3) el = el * S.el14;
  return set;
}

template <typename R, int CAP = FC_CAP, bool PPF = false>
__device__ __forceinline__ int stage_partners(const DevSys& S, int a, const int* __restrict__ row, int p0, int nn,
                                              uint32_t use_sol, uint32_t use_wat, double xr0, double xr1,
                                              double xr2, int* s_j, R (*s_d)[3], int* s_t, int* ngroups,
                                              R (*s_f)[3] = nullptr) {
  const int lane = threadIdx.x % QNB_WAVE;
  const int p = p0 + lane;
  int b = -1, cnt = 0, f0 = 0;
  if (p < nn) {
    b = ldg(row + p);
    if (b < S.ncgp_solute ? use_sol : use_wat) {
      f0 = ldg(S.gfirst + b);
      cnt = ldg(S.glast + b) - f0 + 1;
    }
  }
  int incl = cnt;
#pragma unroll
  for (int o2 = 1; o2 < QNB_WAVE; o2 <<= 1) {
    const int v = __shfl_up(incl, o2, QNB_WAVE);
    if (lane >= o2) incl += v;
  }
  // inclusive sums never decrease, so the groups that fit are a prefix
  const int ng = __popcll(__ballot(incl <= CAP));
  *ngroups = ng;
  if (ng == 0) return 0;
  const int total = __shfl(incl, ng - 1, QNB_WAVE);
  // every group of the batch skipped by the term mask: nothing staged (the
  // slots below hold nothing of this batch)
  if (total == 0) return 0;
  // the group lanes lay out their slots (LDS only): position in cgpatom and
  // the group part of the tag (FT_FAR for a solute group whose atoms are all
  // more than 25 indices from the solute receiver's, neither group having
  // long-range exclusion / 1-4 partners)
  if (lane < ng && cnt > 0) {
    int tag0 = (b < S.ncgp_solute ? FT_SOL : 0) | (owns(a, b) ? FT_OWN : 0);
    if (b < S.ncgp_solute && a < S.ncgp_solute && !ldg(S.glong + a) && !ldg(S.glong + b)) {
      const int alo = ldg(S.gspan + 2 * a), ahi = ldg(S.gspan + 2 * a + 1);
      const int blo = ldg(S.gspan + 2 * b), bhi = ldg(S.gspan + 2 * b + 1);
      if (blo - ahi > 25 || alo - bhi > 25) tag0 |= FT_FAR;
    }
    for (int m = 0, slot = incl - cnt; m < cnt; m++, slot++) {
      s_j[slot] = f0 + m;
      s_t[slot] = tag0 | (m & 3);
    }
  }
  __builtin_amdgcn_wave_barrier();
  // every lane then fills slots lane, lane + 64, ...: FC_SPR slots per round,
  // all atom-index loads of the round first, then all position loads, so the
  // loads of a round are in flight together (slots past the end load the
  // receiver's first atom, a valid index, and store nothing)
  constexpr int FC_SPL = (CAP + QNB_WAVE - 1) / QNB_WAVE;
  constexpr int FC_SPR = FC_SPL % 3 == 0 ? 3 : (FC_SPL % 2 == 0 ? 2 : 1);
  static_assert(FC_SPL % FC_SPR == 0, "staging rounds must cover the slots");
  const int pos_pad = ldg(S.gfirst + a);
#pragma unroll
  for (int u0 = 0; u0 < FC_SPL; u0 += FC_SPR) {
    int jj[FC_SPR], tg[FC_SPR], qa[FC_SPR];
    double px[FC_SPR][3];
#pragma unroll
    for (int u = 0; u < FC_SPR; u++) {
      const int k = lane + (u0 + u) * QNB_WAVE;
      const bool in = k < total;
      tg[u] = in ? s_t[k] : 0;
      jj[u] = ldg(S.cgpatom + (in ? s_j[k] : pos_pad));
    }
    R pf[FC_SPR][3];
#pragma unroll
    for (int u = 0; u < FC_SPR; u++) {
      const int j = jj[u];
      px[u][0] = ldg(S.x + 3 * j);
      px[u][1] = ldg(S.x + 3 * j + 1);
      px[u][2] = ldg(S.x + 3 * j + 2);
      qa[u] = ldg(S.iqatom + j);
      if constexpr (PPF) {
        // the solute partner's LJ factors and charge (ppf / ppd rows); the
        // row of the receiver's first atom stands in for a solvent atom
        const int jp = j < S.nat_solute ? j : ldg(S.cgpatom + pos_pad);
        const R* t = qp_tab<R>(S.ppd, S.ppf) + 4 * (size_t)jp;
        pf[u][0] = ldg(t);
        pf[u][1] = ldg(t + 1);
        pf[u][2] = ldg(t + 2);
      }
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int u = 0; u < FC_SPR; u++) {
      const int k = lane + (u0 + u) * QNB_WAVE;
      if (k < total) {
        s_j[k] = jj[u];
        s_d[k][0] = (R)(px[u][0] - xr0);
        s_d[k][1] = (R)(px[u][1] - xr1);
        s_d[k][2] = (R)(px[u][2] - xr2);
        s_t[k] = tg[u] | (((tg[u] & FT_SOL) && qa[u] != 0) ? FT_Q : 0);
        if constexpr (PPF) {
          s_f[k][0] = pf[u][0];
          s_f[k][1] = pf[u][1];
          s_f[k][2] = pf[u][2];
        }
      }
    }
  }
  __builtin_amdgcn_wave_barrier();
  return total;
}

// ---------------------------------------------------------------------------
// Solute receivers: one wavefront per solute charge group a, its atoms taken
// in sweeps of up to CH (receiver data in LDS and registers).  The row is
// sorted, so the partner solute groups come first and the waters after them.
// Per sweep:
//  * partner solute groups through the staged atom expansion, one partner
//    atom per lane against the sweep's receiver atoms (nonbond_pp,
//    nonbondene.f90:4694-4760).  A group that can hold no excluded or 1-4
//    pair with a (FT_FAR) takes A, B and qq as products of per-atom factors
//    when the system uses the geometric rule with LJ code 1 throughout
//    (ppfast): the products pp_param forms for such a pair, same operands;
//    every other pair goes through pp_param (exclusions, 1-4, ljcod, rule);
//  * partner waters with the fixed mapping of k_wforces (lane l takes site
//    l % 3 of water l / 3 of the round), the pw parameters of that site for
//    every receiver atom held in registers (nonbond_pw 4856-4916);
//  * the pairs inside the group (ig == jgr, i < j, nonbondene.f90:1610) and
//    the receiver side of the Q-atom - solute pairs (nonbond_qp 5157-5230).
// Writes d of the group's atoms, and the wave's pp, pw and qp energy
// partials on the sampled steps (a pp pair's energy on its owner side).
template <typename R, int NS, bool E, int CAP>
__global__ __launch_bounds__(64 * FC_WPB) FC_OCC_ATTR void k_sforces(const DevSys* __restrict__ systems, uint32_t mask,
                                                                     int imap, int nb, int ninst) {
  int inst, blk;
  if (!map_block(imap, nb, ninst, inst, blk)) return;
  const DevSys& S = systems[inst];
  const int wib = threadIdx.x / QNB_WAVE;
  const int a = __builtin_amdgcn_readfirstlane(blk * (int)(blockDim.x / QNB_WAVE) + wib);
  // LJ factor and ljcod tables in LDS (every thread of the block takes part
  // before any wave leaves)
  __shared__ double s_avdw[3 * FC_MAXT], s_bvdw[3 * FC_MAXT];
  __shared__ uint8_t s_ljc[FC_MAXT * FC_MAXT];
  const bool lds_types = S.ntypes <= FC_MAXT;
  if (lds_types) {
    const int nt = S.ntypes;
    for (int k = threadIdx.x; k < 3 * nt; k += blockDim.x) {
      s_avdw[k] = ldg(S.avdw + k);
      s_bvdw[k] = ldg(S.bvdw + k);
    }
    for (int k = threadIdx.x; k < nt * nt; k += blockDim.x) s_ljc[k] = (uint8_t)ldg(S.ljcod + k);
  }
  __syncthreads();
  if (a >= S.ncgp_solute) return;
  const int lane = threadIdx.x % QNB_WAVE;
  extern __shared__ __align__(16) unsigned char fc_smem[];
  unsigned char* wbase = fc_smem + (size_t)wib * sf_wave_bytes<R, CAP>();
  R(*s_pd)[3] = reinterpret_cast<R(*)[3]>(wbase);
  R(*s_o)[3] = reinterpret_cast<R(*)[3]>(wbase + sizeof(R) * 3 * CAP);
  R(*s_pw)[3][3] = reinterpret_cast<R(*)[3][3]>(wbase + sizeof(R) * 3 * (CAP + FC_CH));
  R(*s_pf)[3] = reinterpret_cast<R(*)[3]>(wbase + sizeof(R) * (3 * (CAP + FC_CH) + 9 * FC_CH));
  double* s_rq = reinterpret_cast<double*>(wbase + sizeof(R) * (3 * (2 * CAP + FC_CH) + 9 * FC_CH));
  int* s_pj = reinterpret_cast<int*>(s_rq + FC_CH);
  int* s_pt = s_pj + CAP;
  int* s_ia = s_pt + CAP;
  int* s_rt = s_ia + FC_CH;
  uint32_t* s_rx = reinterpret_cast<uint32_t*>(s_rt + FC_CH);
  uint32_t* s_rm = s_rx + FC_CH;
  const double* __restrict__ x = S.x;
  const auto d = gpm(S.d);
  const int nn = min(ldg(S.nnbr + a), S.cap);
  const int* __restrict__ row = S.nbr + (size_t)a * S.cap;
  const bool want_e = E && want_energies(S);
  const int nsol = S.nat_solute, sa = S.sa, ncs = S.ncgp_solute;
  const int ra = ldg(S.gswitch + a);
  const bool aexcl = ldg(S.excl + ra);
  const bool fast = S.ppfast != 0;
  double e_ppel = 0.0, e_ppvdw = 0.0, e_pwel = 0.0, e_pwvdw = 0.0;
  double eq[2 * NS];  // per state: qp_el, qp_vdw
#pragma unroll
  for (int k = 0; k < 2 * NS; k++) eq[k] = 0.0;
  const double xr0 = ldg(x + (3 * ra)), xr1 = ldg(x + (3 * ra + 1)), xr2 = ldg(x + (3 * ra + 2));
  const int g0 = ldg(S.gfirst + a), na = ldg(S.glast + a) - g0 + 1;
  const int nq = S.nqat, ns = S.nstates;
  // length of the row's solute prefix
  int nsr = 0;
  for (int p0 = 0; p0 < nn; p0 += QNB_WAVE) {
    const int p = p0 + lane;
    const uint64_t m = __ballot(p < nn && ldg(row + p) < ncs);
    nsr += __popcll(m);
    if (m != ~0ull) break;
  }
  const bool qin = (mask & T_QP) && nq > 0 && ldg(S.qp_in + a);
  constexpr int CH = sizeof(R) == 4 ? FC_CH : FC_CH / 2;
  for (int c0 = 0; c0 < na; c0 += CH) {
    const int nk = min(CH, na - c0);
    // receiver atoms of this sweep in LDS: index (-1: takes no pair terms)
    // and offset from the reference atom
    if (lane < nk) {
      const int i = ldg(S.cgpatom + (g0 + c0 + lane));
      s_ia[lane] = (!aexcl && ldg(S.iqatom + i) == 0) ? i : -1;
      s_o[lane][0] = (R)(ldg(x + (3 * i)) - xr0);
      s_o[lane][1] = (R)(ldg(x + (3 * i + 1)) - xr1);
      s_o[lane][2] = (R)(ldg(x + (3 * i + 2)) - xr2);
      s_rt[lane] = ldg(S.iac + i);
      s_rx[lane] = ldg(S.exmask + i);
      s_rm[lane] = ldg(S.m14mask + i);
      s_rq[lane] = ldg(S.crg + i);
      // the atom's solute-water parameters against the 3 water sites
      if (S.nwat > 0 && sa == 3)
        for (int kb = 0; kb < 3; kb++) {
          const size_t tt = (size_t)i + (size_t)nsol * kb;
          s_pw[lane][kb][0] = (R)ldg(S.pwA + tt);
          s_pw[lane][kb][1] = (R)ldg(S.pwB + tt);
          s_pw[lane][kb][2] = (R)ldg(S.pwEl + tt);
        }
    }
    __builtin_amdgcn_wave_barrier();
    // the sweep's receiver data in registers: usable flag, offset, per-atom
    // LJ factors and charge of the fast path
    bool rok[CH];
    R ro[CH][3], rA[CH], rB[CH], rQ[CH];
#pragma unroll
    for (int k = 0; k < CH; k++) {
      const int i = k < nk ? s_ia[k] : -1;
      rok[k] = i >= 0;
      ro[k][0] = s_o[k][0];
      ro[k][1] = s_o[k][1];
      ro[k][2] = s_o[k][2];
      if constexpr (sizeof(R) == 4) {
        rA[k] = rok[k] ? ldg(S.ppf + 4 * i) : 0.0f;
        rB[k] = rok[k] ? ldg(S.ppf + 4 * i + 1) : 0.0f;
        rQ[k] = rok[k] ? ldg(S.ppf + 4 * i + 2) : 0.0f;
      } else {
        rA[k] = rok[k] ? ldg(S.ppd + 4 * i) : 0.0;
        rB[k] = rok[k] ? ldg(S.ppd + 4 * i + 1) : 0.0;
        rQ[k] = rok[k] ? ldg(S.ppd + 4 * i + 2) : 0.0;
      }
    }
    R f[CH][3];
#pragma unroll
    for (int k = 0; k < CH; k++) f[k][0] = f[k][1] = f[k][2] = R(0);
    if (!aexcl) {
      // ---- partner solute groups, one staged atom per lane
      if (mask & T_PP) {
        for (int p0 = 0, ng = 0; p0 < nsr; p0 += ng) {
          const int total = stage_partners<R, CAP, true>(S, a, row, p0, nsr, 1u, 0u, xr0, xr1, xr2, s_pj, s_pd, s_pt,
                                                         &ng, s_pf);
          if (ng == 0) break;
          for (int t = lane; t < total; t += QNB_WAVE) {
            const int tag = s_pt[t], j = s_pj[t];
            if (tag & FT_Q) continue;
            const bool acc = want_e && (tag & FT_OWN);
            const R dx = s_pd[t][0], dy = s_pd[t][1], dz = s_pd[t][2];
            R pel = R(0), pvdw = R(0);
            if (fast && (tag & FT_FAR)) {
              // no exclusion or 1-4 pair possible: A, B, qq from the factors
              // (staged with the partner)
              const R pA = s_pf[t][0], pB = s_pf[t][1], pQ = s_pf[t][2];
#pragma unroll
              for (int k = 0; k < CH; k++) {
                if (!rok[k]) continue;
                const R A = rA[k] * pA, B = rB[k] * pB, el = rQ[k] * pQ;
                // pp_param's `set` test (QREAL_EPS, src/sizes.f90:50); in FP32
                // a pair of zero parameters adds exactly zero, so it is skipped
                // only on the FP64 path
                if constexpr (sizeof(R) == 8)
                  if (!(fabs(A) > R(1.0e-10) || fabs(B) > R(1.0e-10) || fabs(el) > R(1.0e-10))) continue;
                const R ux = dx - ro[k][0], uy = dy - ro[k][1], uz = dz - ro[k][2];
                const R dv = PairMath<R>::nbe(ux, uy, uz, A, B, el, pel, pvdw, acc);
                f[k][0] += ux * dv;
                f[k][1] += uy * dv;
                f[k][2] += uz * dv;
              }
            } else {
              PPAtom Q{j, 0, 0u, 0u, 0.0};
              if (lds_types) {
                Q.t = ldg(S.iac + j);
                Q.xm = ldg(S.exmask + j);
                Q.m14 = ldg(S.m14mask + j);
                Q.q = ldg(S.crg + j);
              }
#pragma unroll
              for (int k = 0; k < CH; k++) {
                if (!rok[k]) continue;
                const int i = s_ia[k];
                double Ad, Bd, eld;
                bool ok;
                if (lds_types) {
                  const PPAtom P{i, s_rt[k], s_rx[k], s_rm[k], s_rq[k]};
                  ok = pp_param_fast(S, P, Q, s_avdw, s_bvdw, s_ljc, Ad, Bd, eld);
                } else {
                  ok = pp_param(S, i, j, Ad, Bd, eld);
                }
                if (!ok) continue;
                const R ux = dx - ro[k][0], uy = dy - ro[k][1], uz = dz - ro[k][2];
                const R dv = PairMath<R>::nbe(ux, uy, uz, (R)Ad, (R)Bd, (R)eld, pel, pvdw, acc);
                f[k][0] += ux * dv;
                f[k][1] += uy * dv;
                f[k][2] += uz * dv;
              }
            }
            if (acc) {
              e_ppel += (double)pel;
              e_ppvdw += (double)pvdw;
            }
          }
          __builtin_amdgcn_wave_barrier();
        }
      }
      // ---- partner waters: lane l takes site kb = l % 3 of water l / 3 of the round
      if ((mask & T_PW) && nsr < nn && sa == 3 && lane < 3 * (QNB_WAVE / 3)) {
        const int kb = lane % 3;
        R pw[CH][3];
#pragma unroll
        for (int k = 0; k < CH; k++) {
          pw[k][0] = rok[k] ? s_pw[k][kb][0] : R(0);
          pw[k][1] = rok[k] ? s_pw[k][kb][1] : R(0);
          pw[k][2] = rok[k] ? s_pw[k][kb][2] : R(0);
        }
        for (int p = nsr + lane / 3; p < nn; p += QNB_WAVE / 3) {
          const int b = ldg(row + p);
          const int j = nsol + sa * (b - ncs) + kb;
          const R dx = (R)(ldg(x + 3 * j) - xr0), dy = (R)(ldg(x + 3 * j + 1) - xr1),
                  dz = (R)(ldg(x + 3 * j + 2) - xr2);
          R pel = R(0), pvdw = R(0);
#pragma unroll
          for (int k = 0; k < CH; k++) {
            if (!rok[k]) continue;
            const R ux = dx - ro[k][0], uy = dy - ro[k][1], uz = dz - ro[k][2];
            const R dv = PairMath<R>::nbe(ux, uy, uz, pw[k][0], pw[k][1], pw[k][2], pel, pvdw, want_e);
            f[k][0] += ux * dv;
            f[k][1] += uy * dv;
            f[k][2] += uz * dv;
          }
          if (want_e) {
            e_pwel += (double)pel;
            e_pwvdw += (double)pvdw;
          }
        }
      }
      // ---- partner molecules of any other solvent: lane l takes site
      // kb = l % sa of molecule l / sa of the round, 64 / sa molecules per
      // round, the parameters read per receiver atom
      if ((mask & T_PW) && nsr < nn && sa != 3 && lane < sa * (QNB_WAVE / sa)) {
        const int kb = lane % sa, M = QNB_WAVE / sa;
        for (int p = nsr + lane / sa; p < nn; p += M) {
          const int b = ldg(row + p);
          const int j = nsol + sa * (b - ncs) + kb;
          const R dx = (R)(ldg(x + 3 * j) - xr0), dy = (R)(ldg(x + 3 * j + 1) - xr1),
                  dz = (R)(ldg(x + 3 * j + 2) - xr2);
          R pel = R(0), pvdw = R(0);
#pragma unroll
          for (int k = 0; k < CH; k++) {
            if (!rok[k]) continue;
            const size_t tt = (size_t)s_ia[k] + (size_t)nsol * kb;
            const R ux = dx - ro[k][0], uy = dy - ro[k][1], uz = dz - ro[k][2];
            const R dv = PairMath<R>::nbe(ux, uy, uz, (R)ldg(S.pwA + tt), (R)ldg(S.pwB + tt), (R)ldg(S.pwEl + tt),
                                          pel, pvdw, want_e);
            f[k][0] += ux * dv;
            f[k][1] += uy * dv;
            f[k][2] += uz * dv;
          }
          if (want_e) {
            e_pwel += (double)pel;
            e_pwvdw += (double)pvdw;
          }
        }
      }
      __builtin_amdgcn_wave_barrier();
      // ---- pairs inside the group: lane (g, kq) = (lane / nk, lane % nk)
      // takes receiver atom kq against group atoms q = g, g + G, ...
      // (G = 64 / nk); an atom of this sweep has its data in LDS, others
      // (groups larger than one sweep) are read from memory
      if (mask & T_PP) {
        const int G = QNB_WAVE / nk, kq = lane % nk, g = lane / nk;
        const int i = s_ia[kq];
        R fq[3] = {R(0), R(0), R(0)};
        if (g < G && i >= 0) {
          const R ox = s_o[kq][0], oy = s_o[kq][1], oz = s_o[kq][2];
          for (int q = g; q < na; q += G) {
            const int ql = q - c0;
            const bool in_sweep = ql >= 0 && ql < nk;
            int j;
            R dx, dy, dz;
            if (in_sweep) {
              j = s_ia[ql];  // -1: a Q-atom, whose pairs belong to the Q kernels
              dx = s_o[ql][0];
              dy = s_o[ql][1];
              dz = s_o[ql][2];
            } else {
              j = ldg(S.cgpatom + (g0 + q));
              if (ldg(S.iqatom + j) != 0) continue;
              dx = (R)(ldg(x + 3 * j) - xr0);
              dy = (R)(ldg(x + 3 * j + 1) - xr1);
              dz = (R)(ldg(x + 3 * j + 2) - xr2);
            }
            if (j < 0 || j == i) continue;
            double A, B, el;
            bool ok;
            if (in_sweep && lds_types) {
              const PPAtom P{i, s_rt[kq], s_rx[kq], s_rm[kq], s_rq[kq]};
              const PPAtom Q{j, s_rt[ql], s_rx[ql], s_rm[ql], s_rq[ql]};
              ok = pp_param_fast(S, P, Q, s_avdw, s_bvdw, s_ljc, A, B, el);
            } else {
              ok = pp_param(S, i, j, A, B, el);
            }
            if (!ok) continue;
            const R ux = dx - ox, uy = dy - oy, uz = dz - oz;
            const R dv = PairMath<R>::nbe(ux, uy, uz, (R)A, (R)B, (R)el, e_ppel, e_ppvdw, want_e && i < j);
            fq[0] += ux * dv;
            fq[1] += uy * dv;
            fq[2] += uz * dv;
          }
        }
#pragma unroll
        for (int k = 0; k < CH; k++)
          if (k == kq) {
            f[k][0] += fq[0];
            f[k][1] += fq[1];
            f[k][2] += fq[2];
          }
      }
    }
    // ---- Q-atom - solute atom, receiver side: lane (g, kq) = (lane / nk,
    // lane % nk) takes receiver atom kq against Q-atoms g, g + G, ...
    if (qin) {
      const R *qtA = qp_tab<R>(S.qpA, S.qpA_f), *qtB = qp_tab<R>(S.qpB, S.qpB_f);
      const R *qtE = qp_tab<R>(S.qpEl, S.qpEl_f), *qtS = qp_tab<R>(S.qpSc, S.qpSc_f);
      const int G = QNB_WAVE / nk, kq = lane % nk, g = lane / nk;
      const int i = s_ia[kq];
      R fq[3] = {R(0), R(0), R(0)};
      if (g < G && i >= 0 && !ldg(S.qconn3 + i)) {
        const R ox = s_o[kq][0], oy = s_o[kq][1], oz = s_o[kq][2];
        for (int iq = g; iq < nq; iq += G) {
          const int q = ldg(S.iqseq + iq);
          const R ux = (R)(ldg(x + 3 * q) - xr0) - ox, uy = (R)(ldg(x + 3 * q + 1) - xr1) - oy,
                  uz = (R)(ldg(x + 3 * q + 2) - xr2) - oz;
          R r2, r;
          rinv(ux * ux + uy * uy + uz * uz, r2, r);
          const R r6 = r2 * r2 * r2;
          R dvs = R(0);
#pragma unroll
          for (int s = 0; s < NS; s++) {
            if (s >= ns) break;
            const size_t t = (size_t)i + (size_t)nsol * ((size_t)iq + (size_t)nq * s);
            dvs += PairMath<R>::nbe_qx(r2, r, r6, ldg(qtA + t), ldg(qtB + t), ldg(qtE + t), ldg(qtS + t),
                                       (R)S.lambda[s], eq[2 * s], eq[2 * s + 1], want_e);
          }
          fq[0] += ux * dvs;
          fq[1] += uy * dvs;
          fq[2] += uz * dvs;
        }
      }
#pragma unroll
      for (int k = 0; k < CH; k++)
        if (k == kq) {
          f[k][0] += fq[0];
          f[k][1] += fq[1];
          f[k][2] += fq[2];
        }
    }
    // reduce all 3 CH force components at once (the per-atom LRF Taylor term
    // follows in k_atom_terms); component c of sweep atom k is value 3k + c
    {
      R fv[3 * CH];
#pragma unroll
      for (int k = 0; k < CH; k++)
#pragma unroll
        for (int c = 0; c < 3; c++) fv[3 * k + c] = f[k][c];
      const double fs = wave_sum_tr(fv, lane);
      const int j = wave_sum_tr_index<3 * CH>(lane);
      if (wave_sum_tr_writer<3 * CH>(lane) && j / 3 < nk) {
        const int i = ldg(S.cgpatom + (g0 + c0 + j / 3));
        d[3 * i + j % 3] = -fs;
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  if (!want_e) return;
  // pp and pw energies, then each state's qp el and vdW
  double part[4 + 2 * NS];
  int slot[4 + 2 * NS];
  part[0] = e_ppel, slot[0] = E_PP_EL;
  part[1] = e_ppvdw, slot[1] = E_PP_VDW;
  part[2] = e_pwel, slot[2] = E_PW_EL;
  part[3] = e_pwvdw, slot[3] = E_PW_VDW;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    part[4 + 2 * s] = eq[2 * s], slot[4 + 2 * s] = E_Q0 + 6 * s + 2;
    part[5 + 2 * s] = eq[2 * s + 1], slot[5 + 2 * s] = E_Q0 + 6 * s + 3;
  }
  store_energy_record(S.epart + (size_t)a * QNB_NE, part, slot, lane);
}

// ---------------------------------------------------------------------------
// Water receivers (rigid 3-site solvent): one wavefront per water a.  Its
// neighbour row is sorted, so the solute groups (indices below ncgp_solute)
// come first, then the waters.  Solute partners go through the staged atom
// expansion, one partner atom per lane against the three receiver atoms with
// the pw parameters (nonbond_pw, nonbondene.f90:4856-4916; energy counted on
// the solute side).  Partner waters need no staging: their atoms are
// consecutive, so in each round lane l takes atom l % 3 of the row's water
// (l / 3) of the round, 21 waters per round (lane 63 idle), reading the
// position straight from x, and evaluates it against the three receiver
// atoms (nonbond_ww_spc nonbondene.f90:5886-5975: O-O LJ + Coulomb, the
// other pairs Coulomb only; nonbond_ww 5509-5573 for other solvents).  Then
// the receiver side of the Q-atom - water pairs.  Writes d of the three
// atoms, and the wave's ww and qw energy partials on the sampled steps.
template <typename R, int NS, bool E, int CAP>
__global__ __launch_bounds__(64 * FC_WPB) void k_wforces(const DevSys* __restrict__ systems, uint32_t mask,
                                                         int imap, int nb, int ninst) {
  int inst, blk;
  if (!map_block(imap, nb, ninst, inst, blk)) return;
  const DevSys& S = systems[inst];
  const int wib = threadIdx.x / QNB_WAVE;
  const int a = __builtin_amdgcn_readfirstlane(S.ncgp_solute + blk * (int)(blockDim.x / QNB_WAVE) + wib);
  if (a >= S.ncgp) return;
  const int lane = threadIdx.x % QNB_WAVE;
  extern __shared__ __align__(16) unsigned char wf_smem[];
  unsigned char* wbase = wf_smem + (size_t)wib * wf_wave_bytes<R, CAP>();
  R(*s_pd)[3] = reinterpret_cast<R(*)[3]>(wbase);
  int* s_pj = reinterpret_cast<int*>(wbase + sizeof(R) * 3 * CAP);
  int* s_pt = s_pj + CAP;
  const double* __restrict__ x = S.x;
  const auto d = gpm(S.d);
  const int nn = min(ldg(S.nnbr + a), S.cap);
  const int* __restrict__ row = S.nbr + (size_t)a * S.cap;
  const bool want_e = E && want_energies(S);
  const int nsol = S.nat_solute, sa = S.sa, ncs = S.ncgp_solute;
  const int ra = nsol + sa * (a - ncs);
  const bool aexcl = ldg(S.excl + ra);
  const int nq = S.nqat, ns = S.nstates;
  double e_el = 0.0, e_vdw = 0.0;
  double eq[2 * NS];  // per state: qw_el, qw_vdw
#pragma unroll
  for (int k = 0; k < 2 * NS; k++) eq[k] = 0.0;
  const double xr0 = ldg(x + 3 * ra), xr1 = ldg(x + 3 * ra + 1), xr2 = ldg(x + 3 * ra + 2);
  R o[3][3], f[3][3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    o[k][0] = (R)(ldg(x + (3 * (ra + k))) - xr0);
    o[k][1] = (R)(ldg(x + (3 * (ra + k) + 1)) - xr1);
    o[k][2] = (R)(ldg(x + (3 * (ra + k) + 2)) - xr2);
    f[k][0] = f[k][1] = f[k][2] = R(0);
  }
  if (!aexcl) {
    // length of the row's solute prefix
    int nsr = 0;
    for (int p0 = 0; p0 < nn; p0 += QNB_WAVE) {
      const int p = p0 + lane;
      const uint64_t m = __ballot(p < nn && ldg(row + p) < ncs);
      nsr += __popcll(m);
      if (m != ~0ull) break;
    }
    // solute partners, one staged atom per lane
    if (mask & T_PW) {
      for (int p0 = 0, ng = 0; p0 < nsr; p0 += ng) {
        const int total = stage_partners<R, CAP>(S, a, row, p0, nsr, 1u, 0u, xr0, xr1, xr2, s_pj, s_pd, s_pt, &ng);
        if (ng == 0) break;
        for (int t = lane; t < total; t += QNB_WAVE) {
          if (s_pt[t] & FT_Q) continue;
          const int j = s_pj[t];
          const R dx = s_pd[t][0], dy = s_pd[t][1], dz = s_pd[t][2];
          R pw[9];
#pragma unroll
          for (int k = 0; k < 3; k++) {
            if constexpr (sizeof(R) == 4) {
              pw[3 * k] = ldg(S.pw_f + PW_REC * j + 3 * k);
              pw[3 * k + 1] = ldg(S.pw_f + PW_REC * j + 3 * k + 1);
              pw[3 * k + 2] = ldg(S.pw_f + PW_REC * j + 3 * k + 2);
            } else {
              const size_t tt = (size_t)j + (size_t)nsol * k;
              pw[3 * k] = (R)ldg(S.pwA + tt);
              pw[3 * k + 1] = (R)ldg(S.pwB + tt);
              pw[3 * k + 2] = (R)ldg(S.pwEl + tt);
            }
          }
#pragma unroll
          for (int k = 0; k < 3; k++) {
            const R ux = dx - o[k][0], uy = dy - o[k][1], uz = dz - o[k][2];
            R ee = R(0), ev = R(0);
            // sites without LJ: nbe with A = B = 0 is nbe_spc exactly
            const R dv = (k > 0 && S.pw_coul_h)
                             ? PairMath<R>::nbe_spc(ux, uy, uz, pw[3 * k + 2], ee, false)
                             : PairMath<R>::nbe(ux, uy, uz, pw[3 * k], pw[3 * k + 1], pw[3 * k + 2], ee, ev, false);
            f[k][0] += ux * dv;
            f[k][1] += uy * dv;
            f[k][2] += uz * dv;
          }
        }
        __builtin_amdgcn_wave_barrier();
      }
    }
    // partner waters: lane l takes atom kb = l % 3 of water (l / 3) of the round
    if (mask & T_WW) {
      R wel[3], wA = R(0), wB = R(0), gA[3], gB[3];
      const int kb = lane % 3, pw0 = lane / 3;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        wel[k] = (R)ldg(S.wwEl + k + 3 * kb);  // receiver atom k with partner atom kb
        gA[k] = (R)ldg(S.wwA + k + 3 * kb);
        gB[k] = (R)ldg(S.wwB + k + 3 * kb);
      }
      if (S.spc && kb == 0) {
        // SPC: LJ for O-O only; the other lanes take the same path with zero
        // LJ parameters, so the k = 0 pair does not branch on kb
        wA = (R)ldg(S.wwA);
        wB = (R)ldg(S.wwB);
      }
      for (int p = nsr + pw0; p < nn; p += QNB_WAVE / 3) {
        if (lane >= 3 * (QNB_WAVE / 3)) break;
        const int b = ldg(row + p);
        const int j = nsol + sa * (b - ncs) + kb;
        const R dx = (R)(ldg(x + 3 * j) - xr0), dy = (R)(ldg(x + 3 * j + 1) - xr1), dz = (R)(ldg(x + 3 * j + 2) - xr2);
        const bool acc = want_e && owns(a, b);
        R pel = R(0), pvdw = R(0);
#pragma unroll
        for (int k = 0; k < 3; k++) {
          const R ux = dx - o[k][0], uy = dy - o[k][1], uz = dz - o[k][2];
          R dv;
          if (S.spc) {
            if (k == 0) dv = PairMath<R>::nbe(ux, uy, uz, wA, wB, wel[0], pel, pvdw, acc);
            else dv = PairMath<R>::nbe_spc(ux, uy, uz, wel[k], pel, acc);
          } else {
            dv = PairMath<R>::nbe(ux, uy, uz, gA[k], gB[k], wel[k], pel, pvdw, acc);
          }
          f[k][0] += ux * dv;
          f[k][1] += uy * dv;
          f[k][2] += uz * dv;
        }
        if (acc) {
          e_el += (double)pel;
          e_vdw += (double)pvdw;
        }
      }
    }
  }
  // Q-atom - water, receiver side (nonbond_qw_spc nonbondene.f90:5658-5767;
  // nonbond_qw 5339-5418): lane per Q-atom, the water's three sites in turn
  // (an SPC hydrogen takes the Coulomb-only nbe_qspc, nonbonded.f90:226-243)
  if ((mask & T_QW) && nq > 0 && ldg(S.qw_in + (a - ncs))) {
    for (int iq = lane; iq < nq; iq += QNB_WAVE) {
      const int q = ldg(S.iqseq + iq);
      const R qx = (R)(ldg(x + 3 * q) - xr0), qy = (R)(ldg(x + 3 * q + 1) - xr1), qz = (R)(ldg(x + 3 * q + 2) - xr2);
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const R ux = qx - o[k][0], uy = qy - o[k][1], uz = qz - o[k][2];
        R r2, r;
        rinv(ux * ux + uy * uy + uz * uz, r2, r);
        R dvs = R(0);
        if (k == 0 || !S.spc) {
          const R r6 = r2 * r2 * r2;
#pragma unroll
          for (int s = 0; s < NS; s++) {
            if (s >= ns) break;
            const size_t t = (size_t)iq + (size_t)nq * ((size_t)k + (size_t)sa * s);
            dvs += PairMath<R>::nbe_qx(r2, r, r6, (R)ldg(S.qwA + t), (R)ldg(S.qwB + t), (R)ldg(S.qwEl + t),
                                       (R)ldg(S.qwSc + t), (R)S.lambda[s], eq[2 * s], eq[2 * s + 1], want_e);
          }
        } else {
#pragma unroll
          for (int s = 0; s < NS; s++) {
            if (s >= ns) break;
            const size_t t = (size_t)iq + (size_t)nq * ((size_t)k + (size_t)sa * s);
            const R Vel = (R)ldg(S.qwEl + t) * r;
            if (want_e) eq[2 * s] += (double)Vel;
            dvs += -r2 * Vel * (R)S.lambda[s];
          }
        }
        f[k][0] += ux * dvs;
        f[k][1] += uy * dvs;
        f[k][2] += uz * dvs;
      }
    }
  }
  {
    // the nine force components in one transposed reduction; component c of
    // atom k is value 3k + c, i.e. d[3 * ra + j]
    R fv[9];
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int c = 0; c < 3; c++) fv[3 * k + c] = f[k][c];
    const double fs = wave_sum_tr(fv, lane);
    if (wave_sum_tr_writer<9>(lane)) d[3 * ra + wave_sum_tr_index<9>(lane)] = -fs;
  }
  if (!want_e) return;
  // ww energies, then each state's qw el and vdW
  double part[2 + 2 * NS];
  int slot[2 + 2 * NS];
  part[0] = e_el, slot[0] = E_WW_EL;
  part[1] = e_vdw, slot[1] = E_WW_VDW;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    part[2 + 2 * s] = eq[2 * s], slot[2 + 2 * s] = E_Q0 + 6 * s + 4;
    part[3 + 2 * s] = eq[2 * s + 1], slot[3 + 2 * s] = E_Q0 + 6 * s + 5;
  }
  store_energy_record(S.epart + (size_t)a * QNB_NE, part, slot, lane);
}

// ---------------------------------------------------------------------------
// Receivers of a solvent other than 3-site water (any solv_atom up to 64):
// one wavefront per solvent molecule a; lane l takes receiver atom
// k = l % sa and, with M = 64 / sa lanes per atom, every M-th partner atom
// from m = l / sa on, in the partner order of the reference's loops:
//  * solute partner groups (the row's prefix): their non-Q atoms with the pw
//    parameters (nonbond_pw, nonbondene.f90:4856-4916; energy counted on the
//    solute side);
//  * partner molecules: the sa x sa atom pairs with ww_precomp
//    (nonbond_ww 5509-5573), energy on the owner side of the pair;
//  * the Q-atoms (receiver side of nonbond_qw 5339-5418);
//  * the pairs inside the molecule (nonbond_solvent_internal 6082-6154),
//    excluded molecules included: the lane of each end atom evaluates the
//    pair, the first end counts its energy.
// The lanes of an atom add their FP64 partials in lane order (LDS); writes d
// of the molecule's atoms and the wave's ww and qw energy partials on the
// sampled steps.
template <typename R, int NS>
__global__ __launch_bounds__(64 * FC_WPB) void k_gforces(const DevSys* __restrict__ systems, uint32_t mask,
                                                         int imap, int nb, int ninst) {
  int inst, blk;
  if (!map_block(imap, nb, ninst, inst, blk)) return;
  const DevSys& S = systems[inst];
  const int wib = threadIdx.x / QNB_WAVE;
  const int a = __builtin_amdgcn_readfirstlane(S.ncgp_solute + blk * (int)(blockDim.x / QNB_WAVE) + wib);
  if (a >= S.ncgp) return;
  __shared__ double s_f[FC_WPB][QNB_WAVE][3];
  const int lane = threadIdx.x % QNB_WAVE;
  const int sa = S.sa, M = QNB_WAVE / sa;
  const int k = lane % sa, m = lane / sa;
  const bool active = m < M;
  const double* __restrict__ x = S.x;
  const int nsol = S.nat_solute, ncs = S.ncgp_solute, nq = S.nqat, ns = S.nstates;
  const int base = nsol + sa * (a - ncs), ia = base + k;
  const bool aexcl = ldg(S.excl + base);
  const bool want_e = want_energies(S);
  const int nn = min(ldg(S.nnbr + a), S.cap);
  const int* __restrict__ row = S.nbr + (size_t)a * S.cap;
  const double xi0 = ldg(x + 3 * ia), xi1 = ldg(x + 3 * ia + 1), xi2 = ldg(x + 3 * ia + 2);
  double f[3] = {0.0, 0.0, 0.0};
  double e_el = 0.0, e_vdw = 0.0;
  double eq[2 * NS];
#pragma unroll
  for (int c = 0; c < 2 * NS; c++) eq[c] = 0.0;
  if (active && !aexcl) {
    for (int p = 0; p < nn; p++) {
      const int b = ldg(row + p);
      if (b < ncs) {
        if (!(mask & T_PW)) continue;
        const int t1 = ldg(S.glast + b);
        for (int t = ldg(S.gfirst + b) + m; t <= t1; t += M) {
          const int j = ldg(S.cgpatom + t);
          if (ldg(S.iqatom + j) != 0) continue;
          const R ux = (R)(ldg(x + 3 * j) - xi0), uy = (R)(ldg(x + 3 * j + 1) - xi1),
                  uz = (R)(ldg(x + 3 * j + 2) - xi2);
          const size_t tt = (size_t)j + (size_t)nsol * k;
          R ee = R(0), ev = R(0);
          const R dv = PairMath<R>::nbe(ux, uy, uz, (R)ldg(S.pwA + tt), (R)ldg(S.pwB + tt), (R)ldg(S.pwEl + tt),
                                        ee, ev, false);
          f[0] += (double)(ux * dv);
          f[1] += (double)(uy * dv);
          f[2] += (double)(uz * dv);
        }
      } else {
        if (!(mask & T_WW)) continue;
        const bool own = want_e && owns(a, b);
        const int jb = nsol + sa * (b - ncs);
        for (int kb = m; kb < sa; kb += M) {
          const int j = jb + kb;
          const R ux = (R)(ldg(x + 3 * j) - xi0), uy = (R)(ldg(x + 3 * j + 1) - xi1),
                  uz = (R)(ldg(x + 3 * j + 2) - xi2);
          const int t = k + kb * sa;
          R ee = R(0), ev = R(0);
          const R dv = (S.spc && (k != 0 || kb != 0))
                           ? PairMath<R>::nbe_spc(ux, uy, uz, (R)ldg(S.wwEl + t), ee, own)
                           : PairMath<R>::nbe(ux, uy, uz, (R)ldg(S.wwA + t), (R)ldg(S.wwB + t), (R)ldg(S.wwEl + t),
                                              ee, ev, own);
          f[0] += (double)(ux * dv);
          f[1] += (double)(uy * dv);
          f[2] += (double)(uz * dv);
          if (own) {
            e_el += (double)ee;
            e_vdw += (double)ev;
          }
        }
      }
    }
    if ((mask & T_QW) && nq > 0 && ldg(S.qw_in + (a - ncs))) {
      for (int iq = m; iq < nq; iq += M) {
        const int q = ldg(S.iqseq + iq);
        const R ux = (R)(ldg(x + 3 * q) - xi0), uy = (R)(ldg(x + 3 * q + 1) - xi1), uz = (R)(ldg(x + 3 * q + 2) - xi2);
        R r2, r;
        rinv(ux * ux + uy * uy + uz * uz, r2, r);
        R dvs = R(0);
        if (k == 0 || !S.spc) {
          const R r6 = r2 * r2 * r2;
#pragma unroll
          for (int st = 0; st < NS; st++) {
            if (st >= ns) break;
            const size_t t = (size_t)iq + (size_t)nq * ((size_t)k + (size_t)sa * st);
            dvs += PairMath<R>::nbe_qx(r2, r, r6, (R)ldg(S.qwA + t), (R)ldg(S.qwB + t), (R)ldg(S.qwEl + t),
                                       (R)ldg(S.qwSc + t), (R)S.lambda[st], eq[2 * st], eq[2 * st + 1], want_e);
          }
        } else {
#pragma unroll
          for (int st = 0; st < NS; st++) {
            if (st >= ns) break;
            const size_t t = (size_t)iq + (size_t)nq * ((size_t)k + (size_t)sa * st);
            const R Vel = (R)ldg(S.qwEl + t) * r;
            if (want_e) eq[2 * st] += (double)Vel;
            dvs += -r2 * Vel * (R)S.lambda[st];
          }
        }
        f[0] += (double)(ux * dvs);
        f[1] += (double)(uy * dvs);
        f[2] += (double)(uz * dvs);
      }
    }
  }
  if ((mask & T_SINT) && S.n_sint > 0 && active && m == 0) {
    for (int ip = 0; ip < S.n_sint; ip++) {
      const int oa = ldg(S.sint_ij + 2 * ip), ob = ldg(S.sint_ij + 2 * ip + 1);
      if (oa != k && ob != k) continue;
      const int i0 = base + oa, i1 = base + ob;
      const R vx = (R)(ldg(x + 3 * i1) - ldg(x + 3 * i0)), vy = (R)(ldg(x + 3 * i1 + 1) - ldg(x + 3 * i0 + 1)),
              vz = (R)(ldg(x + 3 * i1 + 2) - ldg(x + 3 * i0 + 2));
      R r2, r;
      rinv(vx * vx + vy * vy + vz * vz, r2, r);
      const R r6 = r2 * r2 * r2, r12 = r6 * r6;
      const R Vel = (R)ldg(S.sint_par + 3 * ip + 2) * r, Va = (R)ldg(S.sint_par + 3 * ip) * r12,
              Vb = (R)ldg(S.sint_par + 3 * ip + 1) * r6;
      const R dv = r2 * (-Vel) + r2 * (R(-12) * Va + R(6) * Vb);
      // d(ia) -= v dv, d(ib) += v dv; the lane holds the negative gradient
      // part f (d = -f), as for the pairs above (u = partner - receiver)
      const double sgn = oa == k ? 1.0 : -1.0;
      if (oa == k && ob == k) continue;  // a self pair carries no force
      f[0] += sgn * (double)(vx * dv);
      f[1] += sgn * (double)(vy * dv);
      f[2] += sgn * (double)(vz * dv);
      if (want_e && oa == k) {
        e_el += (double)Vel;
        e_vdw += (double)(Va - Vb);
      }
    }
  }
  s_f[wib][lane][0] = f[0];
  s_f[wib][lane][1] = f[1];
  s_f[wib][lane][2] = f[2];
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  if (lane < sa) {
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
    for (int mm = 0; mm < M; mm++) {
      t0 += s_f[wib][lane + sa * mm][0];
      t1 += s_f[wib][lane + sa * mm][1];
      t2 += s_f[wib][lane + sa * mm][2];
    }
    const auto d = gpm(S.d);
    d[3 * ia] = -t0;
    d[3 * ia + 1] = -t1;
    d[3 * ia + 2] = -t2;
  }
  if (!want_e) return;
  double part[2 + 2 * NS];
  int slot[2 + 2 * NS];
  part[0] = e_el, slot[0] = E_WW_EL;
  part[1] = e_vdw, slot[1] = E_WW_VDW;
#pragma unroll
  for (int st = 0; st < NS; st++) {
    part[2 + 2 * st] = eq[2 * st], slot[2 + 2 * st] = E_Q0 + 6 * st + 4;
    part[3 + 2 * st] = eq[2 * st + 1], slot[3 + 2 * st] = E_Q0 + 6 * st + 5;
  }
  store_energy_record(S.epart + (size_t)a * QNB_NE, part, slot, lane);
}

// ---------------------------------------------------------------------------
// Per-atom terms after the pair kernels, one thread per atom: the LRF Taylor
// term (lrf_taylor, nonbondene.f90:507-571) of every non-Q, non-excluded
// atom and the SCAAS radial restraint (restrain_solvent, 6466-6544) of the
// solvent atoms radial_source names; adds to d (store_atom_gradient).  On the
// sampled steps the atom's two energies go to eatom[2i], eatom[2i+1] for the
// fixed-order reduction.
__global__ __launch_bounds__(256) void k_atom_terms(const DevSys* __restrict__ systems, uint32_t mask,
                                                    const double* __restrict__ tfree) {
  const DevSys& S = systems[blockIdx.y];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= S.natom) return;
  double f[3] = {0.0, 0.0, 0.0};
  double el = 0.0, er = 0.0;
  if ((mask & T_LRF) && S.use_lrf && !ldg(S.excl + i) && ldg(S.iqatom + i) == 0) lrf_taylor_atom(S, i, f, el);
  const int src = (mask & T_RAD) ? radial_source(S, i) : -1;
  if (src >= 0) {
    const double T = tfree[blockIdx.y];
    const double shift = S.fk_wsphere != 0.0 ? sqrt(S.Boltz * T / S.fk_wsphere) : 0.0;
    radial_restraint(S, src, shift, f, er);
  }
  store_atom_gradient(S, i, f);
  if (want_energies(S)) {
    stg(S.eatom + 2 * i, el);
    stg(S.eatom + 2 * i + 1, er);
  }
}

// ---------------------------------------------------------------------------
// Q-atom side of the Q-water and Q-solute pairs: wave = (Q-atom, chunk of its
// partners, 64 at a time); the wave's sum is added to the Q-atom's gradient
// with atomics (a few per Q-atom), or, with dq, to the system's row
// dq[blockIdx.y * dq_stride + 3 iq ..] (the MD step runs this beside the
// solute receivers, which store the same atoms' rows, and adds dq in its
// per-atom pass).
template <typename R, int NS>
__global__ __launch_bounds__(256) void k_qforces(const DevSys* __restrict__ systems, uint32_t mask,
                                                 int chunks_per_q, double* __restrict__ dq, int dq_stride) {
  const DevSys& S = systems[blockIdx.y];
  // block x runs on XCD x % 8 and takes Q-atom iq = x % 8 + 8 * (x / 8 / bq),
  // chunks (x / 8 % bq) * W .. + W - 1 (W waves per block, bq blocks per
  // Q-atom): an XCD serves only the Q-atoms iq = its index mod 8, so the rows
  // of the Q tables it reads stay in its L2 across the instances
  const int W = blockDim.x / QNB_WAVE, bq = (chunks_per_q + W - 1) / W;
  const int m = blockIdx.x / 8;
  const int iq = blockIdx.x % 8 + 8 * (m / bq), ch = (m % bq) * W + threadIdx.x / QNB_WAVE;
  const int lane = threadIdx.x % QNB_WAVE;
  if (iq >= S.nqat || ch >= chunks_per_q) return;
  const int nqw = ldg(S.nq_lists + 0), nqp = ldg(S.nq_lists + 1);
  const int nw_ch = (mask & T_QW) ? (nqw + QNB_WAVE - 1) / QNB_WAVE : 0;
  const int np_ch = (mask & T_QP) ? (nqp + QNB_WAVE - 1) / QNB_WAVE : 0;
  const double* __restrict__ x = S.x;
  const int i = ldg(S.iqseq + iq);
  const int nq = S.nqat, ns = S.nstates, sa = S.sa;
  const double xi0 = ldg(x + 3 * i), xi1 = ldg(x + 3 * i + 1), xi2 = ldg(x + 3 * i + 2);
  // per-lane partials in R (a lane takes one partner per chunk), summed over
  // the wave in FP64
  R f[3] = {R(0), R(0), R(0)};
  double dummy_e = 0.0, dummy_v = 0.0;
  const R *qtA = qp_tab<R>(S.qpA, S.qpA_f), *qtB = qp_tab<R>(S.qpB, S.qpB_f);
  const R *qtE = qp_tab<R>(S.qpEl, S.qpEl_f), *qtS = qp_tab<R>(S.qpSc, S.qpSc_f);
  for (int c = ch; c < nw_ch + np_ch; c += chunks_per_q) {
    const int p = c * QNB_WAVE + lane;
    if (c < nw_ch) {
      if (p >= nqw) continue;
      const int o = S.nat_solute + sa * ldg(S.qw_list + p);
      for (int k = 0; k < sa; k++) {
        const int j = o + k;
        const double ux = ldg(x + 3 * j) - xi0, uy = ldg(x + 3 * j + 1) - xi1, uz = ldg(x + 3 * j + 2) - xi2;
        const R rux = (R)ux, ruy = (R)uy, ruz = (R)uz;
        R r2, r;
        rinv(rux * rux + ruy * ruy + ruz * ruz, r2, r);
        R dvs = R(0);
        if (k == 0 || !S.spc) {
          R r6 = r2 * r2 * r2;
#pragma unroll
          for (int s = 0; s < NS; s++) {
            if (s >= ns) break;
            const size_t t = (size_t)iq + (size_t)nq * ((size_t)k + (size_t)sa * s);
            dvs += PairMath<R>::nbe_qx(r2, r, r6, (R)ldg(S.qwA + t), (R)ldg(S.qwB + t), (R)ldg(S.qwEl + t),
                                       (R)ldg(S.qwSc + t),
                                       (R)S.lambda[s], dummy_e, dummy_v, false);
          }
        } else {
#pragma unroll
          for (int s = 0; s < NS; s++) {
            if (s >= ns) break;
            const size_t t = (size_t)iq + (size_t)nq * ((size_t)k + (size_t)sa * s);
            dvs += -r2 * ((R)ldg(S.qwEl + t) * r) * (R)S.lambda[s];
          }
        }
        f[0] -= rux * dvs;
        f[1] -= ruy * dvs;
        f[2] -= ruz * dvs;
      }
    } else {
      const int pp = p - nw_ch * QNB_WAVE;
      if (pp >= nqp) continue;
      const int j = ldg(S.qp_list + pp);
      const double ux = ldg(x + 3 * j) - xi0, uy = ldg(x + 3 * j + 1) - xi1, uz = ldg(x + 3 * j + 2) - xi2;
      const R rux = (R)ux, ruy = (R)uy, ruz = (R)uz;
      R r2, r;
      rinv(rux * rux + ruy * ruy + ruz * ruz, r2, r);
      R r6 = r2 * r2 * r2;
#pragma unroll
      for (int s = 0; s < NS; s++) {
        if (s >= ns) break;
        const size_t t = (size_t)j + (size_t)S.nat_solute * ((size_t)iq + (size_t)nq * s);
        R dv = PairMath<R>::nbe_qx(r2, r, r6, ldg(qtA + t), ldg(qtB + t), ldg(qtE + t), ldg(qtS + t),
                                   (R)S.lambda[s], dummy_e, dummy_v, false);
        f[0] -= rux * dv;
        f[1] -= ruy * dv;
        f[2] -= ruz * dv;
      }
    }
  }
  const double fs = wave_sum_tr(f, lane);
  const int c = wave_sum_tr_index<3>(lane);
  if (wave_sum_tr_writer<3>(lane) && fs != 0.0)
    unsafeAtomicAdd(dq ? dq + (size_t)blockIdx.y * dq_stride + 3 * iq + c : &S.d[3 * i + c], fs);
}

// ---------------------------------------------------------------------------
// Static Q-Q (nonbond_qq, nonbondene.f90:5013-5082, pair kernel nbe_qq
// nonbonded.f90:111-150) and Q-bonded-neighbour (nonbond_qqp, 5085-5153)
// lists: one block of QQ_THREADS per system, a wavefront per atom of the
// lists (entity) whose lanes evaluate the pairs the atom belongs to; the
// wave's sum goes to the atom's gradient once (no atomics; each pair is
// evaluated from both sides).  A pair's energy is counted by its first atom;
// per state the energies are summed per wave and then over the waves in
// order.
template <int NS>
__device__ __forceinline__ void qq_add(double (&eq)[NS][4], int s, int c0, double v0, double v1) {
#pragma unroll
  for (int k = 0; k < NS; k++)
    if (k == s) {
      eq[k][c0] += v0;
      eq[k][c0 + 1] += v1;
    }
}

template <typename R, int NS>
__global__ __launch_bounds__(QQ_THREADS) void k_qq(const DevSys* __restrict__ systems, uint32_t mask) {
  const DevSys& S = systems[blockIdx.x];
  const int slot = S.ncgp;
  const int t = threadIdx.x, lane = t % QNB_WAVE, wid = t / QNB_WAVE;
  constexpr int NW = QQ_THREADS / QNB_WAVE;
  const auto x = gp(S.x);
  const auto d = gpm(S.d);
  const bool want_e = want_energies(S);
  double eq[NS][4];
#pragma unroll
  for (int s = 0; s < NS; s++) eq[s][0] = eq[s][1] = eq[s][2] = eq[s][3] = 0.0;
  // wave wid takes entities wid, wid + NW, ...; its lanes take the entity's pairs
  for (int e = wid; e < S.nqe; e += NW) {
    const int me = ldg(S.qe_atom + e);
    double f[3] = {0.0, 0.0, 0.0};
    for (int it = ldg(S.qe_off + e) + lane, end = ldg(S.qe_off + e + 1); it < end; it += QNB_WAVE) {
      const int i = ldg(S.qe_item + 4 * it), j = ldg(S.qe_item + 4 * it + 1), p = ldg(S.qe_item + 4 * it + 2);
      const int w = ldg(S.qe_item + 4 * it + 3), s = w >> 2, side = w & 1;
      const bool qq = (w & 2) == 0;
      if (!(mask & (qq ? T_QQ : T_QQP))) continue;
      const double* pr = (qq ? S.qq_par : S.qqp_par) + 4 * p;
      const double ux = x[3 * j] - x[3 * i], uy = x[3 * j + 1] - x[3 * i + 1], uz = x[3 * j + 2] - x[3 * i + 2];
      const R rux = (R)ux, ruy = (R)uy, ruz = (R)uz;
      R r2, r;
      rinv(rux * rux + ruy * ruy + ruz * ruz, r2, r);
      const R pA = (R)ldg(pr), pB = (R)ldg(pr + 1), pel = (R)ldg(pr + 2), psc = (R)ldg(pr + 3);
      const R lam = (R)S.lambda[s];
      R dv;
      if (qq) {
        R r6d = r2 * r2 * r2;
        R r6_hc = recip(r6d);
        R r6 = recip(r6_hc + psc);
        R r12 = r6 * r6;
        R Vel = pel * r, Va, Vb;
        if (ldg(S.qq_soft + p)) {
          Vb = R(0);
          Va = pA * exp(-pB / r);
          dv = r2 * (-Vel - pB * Va / r) * lam;
        } else {
          Va = pA * r12;
          Vb = pB * r6;
          dv = r2 * (-Vel - (R(12) * Va - R(6) * Vb) * r6 * r6_hc) * lam;
        }
        if (want_e && side == 0) qq_add<NS>(eq, s, 0, (double)Vel, (double)Va - (double)Vb);
      } else {
        double el = 0.0, vdw = 0.0;
        dv = PairMath<R>::nbe_qx(r2, r, r2 * r2 * r2, pA, pB, pel, psc, lam, el, vdw, want_e && side == 0);
        if (want_e && side == 0) qq_add<NS>(eq, s, 2, el, vdw);
      }
      // the pair's gradient is -u dv on its first atom, +u dv on the second
      const double sg = side == 0 ? -1.0 : 1.0;
      f[0] += sg * (double)(rux * dv);
      f[1] += sg * (double)(ruy * dv);
      f[2] += sg * (double)(ruz * dv);
    }
    f[0] = wave_sum(f[0]);
    f[1] = wave_sum(f[1]);
    f[2] = wave_sum(f[2]);
    if (lane == 0) {
      d[3 * me] += f[0];
      d[3 * me + 1] += f[1];
      d[3 * me + 2] += f[2];
    }
  }
  if (!want_e) return;
  __shared__ double red[NW][NS][4];
#pragma unroll
  for (int s = 0; s < NS; s++)
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const double v = wave_sum(eq[s][c]);
      if (lane == 0) red[wid][s][c] = v;
    }
  __syncthreads();
  if (t != 0) return;
  double* ep = S.epart + (size_t)slot * QNB_NE;
  for (int k = 0; k < QNB_NE; k++) ep[k] = 0.0;
  for (int s = 0; s < NS; s++)
    for (int c = 0; c < 4; c++) {
      double v = 0.0;
      for (int w = 0; w < NW; w++) v += red[w][s][c];
      ep[E_Q0 + 6 * s + c] = v;
    }
}

// ---------------------------------------------------------------------------
// SCAAS polarisation restraint (watpol, nonbondene.f90:6548-6746,
// md = .true.): one block per system.  The reference's per-shell
// selection sort by theta is a stable ascending sort; here each member's rank
// is counted directly (ties broken by water index, as the selection sort does).
// Dipole and radial directions of the molecule whose first atom is i
// (nonbondene.f90:6572-6604): a 3-site solvent takes H1 + H2 - 2 O and the
// radial vector of O; any other the charge-weighted sum of its positions
// (charges of the first solvent molecule) and its centroid's radial vector.
__device__ __forceinline__ void wat_vectors(const DevSys& S, int i, double* rmu, double& rm, double* rcu,
                                            double& rc) {
  const double* x = S.x;
  const int sa = S.sa;
  double m[3], v[3];
  if (sa == 3) {
    for (int c = 0; c < 3; c++) m[c] = (x[3 * (i + 1) + c] + x[3 * (i + 2) + c]) - x[3 * i + c] * 2.0;
    for (int c = 0; c < 3; c++) v[c] = x[3 * i + c] - S.xw[c];
  } else {
    const double* chg = S.crg + S.nat_solute;
    double g[3] = {0.0, 0.0, 0.0};
    m[0] = m[1] = m[2] = 0.0;
    for (int k = 0; k < sa; k++) {
      const double q = ldg(chg + k);
      for (int c = 0; c < 3; c++) {
        g[c] = g[c] + x[3 * (i + k) + c];
        m[c] = m[c] + x[3 * (i + k) + c] * q;
      }
    }
    for (int c = 0; c < 3; c++) v[c] = g[c] / (double)sa - S.xw[c];
  }
  rm = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
  for (int c = 0; c < 3; c++) rmu[c] = m[c] / rm;
  rc = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int c = 0; c < 3; c++) rcu[c] = v[c] / rc;
}

__device__ __forceinline__ bool has_pol(const DevSys& S) { return S.nwat > 0 && S.wpol && S.nshell > 0; }

// Phase 1, one thread per water: theta and shell (-1: inside the innermost
// shell, or excluded); the members of each shell are appended to its list
// (slot order is immaterial: ranks are counted by value).  Block 0 first
// applies the theta_corr feedback of the steps that are multiples of itdis
// (nonbondene.f90:6653-6666).  The shell counts start at zero (set at
// creation and cleared by phase 4).
__global__ __launch_bounds__(WP_THREADS) void k_watpol_theta(const DevSys* __restrict__ systems,
                                                             const int* __restrict__ istep) {
  const DevSys& S = systems[blockIdx.y];
  if (!has_pol(S)) return;
  const int t = threadIdx.x, nsh = S.nshell, nw = S.nwat;
  if (blockIdx.x == 0 && t == 0) {
    double* st = S.sh_state;  // theta_corr[8], avtheta[8], avn_insh[8]
    const int step = istep[blockIdx.y];
    const int itd = (int)llround(S.itdis);
    if (step > 0 && itd > 0 && step % itd == 0) {
      for (int is = 0; is < nsh; is++) {
        st[QNB_MAXSH + is] = st[QNB_MAXSH + is] / S.itdis;
        st[2 * QNB_MAXSH + is] = st[2 * QNB_MAXSH + is] / S.itdis;
        st[is] = st[is] + st[QNB_MAXSH + is] - acos(S.sh_cstb[is]);
        st[QNB_MAXSH + is] = 0.0;
        st[2 * QNB_MAXSH + is] = 0.0;
      }
    }
  }
  const int w = blockIdx.x * WP_THREADS + t;
  if (w >= nw) return;
  const int i = S.nat_solute + S.sa * w;
  int sh = -1;
  double th = 0.0;
  if (!ldg(S.excl + i)) {
    double rmu[3], rcu[3], rm, rc;
    wat_vectors(S, i, rmu, rm, rcu, rc);
    double scp = rmu[0] * rcu[0] + rmu[1] * rcu[1] + rmu[2] * rcu[2];
    scp = scp > 1.0 ? 1.0 : (scp < -1.0 ? -1.0 : scp);
    th = acos(scp);
    if (rc > S.sh_rout[nsh - 1] - S.sh_dr[nsh - 1]) {
      int is;
      for (is = nsh - 1; is >= 1; is--)
        if (rc <= S.sh_rout[is]) break;
      if (is < 1) is = 0;
      sh = is;
    }
  }
  stg(S.w_theta + w, th);
  stg(S.w_shell + w, sh);
  // append to the shell lists with one atomic per (wave, shell): the lanes of
  // a shell take consecutive places (k_watpol_rank orders each shell by
  // (theta, water index), so the order of appending does not matter)
  const int lane = t % QNB_WAVE;
  int pending = sh;
  for (;;) {
    const uint64_t act = __ballot(pending >= 0);
    if (act == 0) break;
    const int lead = __ffsll((unsigned long long)act) - 1;
    const int s0 = __shfl(pending, lead, QNB_WAVE);
    const uint64_t m = __ballot(pending == s0);
    int base = 0;
    if (lane == lead) base = atomicAdd(S.sh_n + s0, __popcll(m));
    base = __shfl(base, lead, QNB_WAVE);
    if (pending == s0) {
      const int pos = base + __popcll(m & ((1ull << lane) - 1ull));
      stg(S.sh_list + (size_t)s0 * nw + pos, w);
      stg(S.sh_lth + (size_t)s0 * nw + pos, th);
      pending = -1;
    }
  }
}

// Phase 2, one block per (shell, system): the rank of every member in the
// shell's stable ascending order of theta (the reference's selection sort;
// ties by water index).  Shells of up to WP_SORT members are sorted by
// (theta, water index) with a bitonic network in LDS and the rank is the
// sorted position; larger shells count, for every member, the members
// before it (staged through LDS in chunks of WP_SORT).
#define WP_SORT 4096
__global__ __launch_bounds__(WP_RANK_THREADS) void k_watpol_rank(const DevSys* __restrict__ systems) {
  const DevSys& S = systems[blockIdx.y];
  const int sh = blockIdx.x;
  if (!has_pol(S) || sh >= S.nshell) return;
  __shared__ double c_th[WP_SORT];
  __shared__ int c_w[WP_SORT];
  const int t = threadIdx.x, nw = S.nwat;
  const int n = S.sh_n[sh];
  const int* L = S.sh_list + (size_t)sh * nw;
  const double* Lth = S.sh_lth + (size_t)sh * nw;
  if (n <= WP_SORT) {
    int N = 2;
    while (N < n) N <<= 1;
    for (int k = t; k < N; k += WP_RANK_THREADS) {
      c_th[k] = k < n ? ldg(Lth + k) : __builtin_inf();
      c_w[k] = k < n ? ldg(L + k) : 0x7fffffff;
    }
    __syncthreads();
    for (int k = 2; k <= N; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = t; i < N; i += WP_RANK_THREADS) {
          const int l = i ^ j;
          if (l > i) {
            const double ti = c_th[i], tl = c_th[l];
            const int wi = c_w[i], wl = c_w[l];
            const bool gt = ti > tl || (ti == tl && wi > wl);  // (i) after (l)
            if (((i & k) == 0) == gt) {
              c_th[i] = tl;
              c_th[l] = ti;
              c_w[i] = wl;
              c_w[l] = wi;
            }
          }
        }
        __syncthreads();
      }
    for (int p = t; p < n; p += WP_RANK_THREADS) stg(S.w_rank + c_w[p], p);
    return;
  }
  for (int m0 = 0; m0 < n; m0 += WP_RANK_THREADS) {
    const int m = m0 + t;
    const int w = m < n ? ldg(L + m) : -1;
    const double th = m < n ? ldg(Lth + m) : 0.0;
    int r = 0;
    for (int k0 = 0; k0 < n; k0 += WP_SORT) {
      const int nk = min(WP_SORT, n - k0);
      __syncthreads();
      for (int k = t; k < nk; k += WP_RANK_THREADS) {
        c_w[k] = ldg(L + k0 + k);
        c_th[k] = ldg(Lth + k0 + k);
      }
      __syncthreads();
      if (m < n)
        for (int k = 0; k < nk; k++) {
          const double t2 = c_th[k];
          r += (t2 < th) || (t2 == th && c_w[k] < w);
        }
    }
    if (m < n) stg(S.w_rank + w, r);
  }
}

// Phase 3, one thread per water of a shell: target angle from the rank, the
// restraint energy and forces (adds to d), and per-block sums of the energy
// and of theta per shell into wp_part.
__global__ __launch_bounds__(WP_THREADS) void k_watpol_force(const DevSys* __restrict__ systems) {
  const DevSys& S = systems[blockIdx.y];
  const int t = threadIdx.x, lane = t % QNB_WAVE, wid = t / QNB_WAVE;
  if (!has_pol(S)) return;
  const int nsh = S.nshell, nw = S.nwat;
  // blocks of this system (the grid is sized for the batch's largest)
  const int nblk = (nw + WP_THREADS - 1) / WP_THREADS;
  if ((int)blockIdx.x >= nblk) return;
  const double pi = 3.141592653589793;
  const int w = blockIdx.x * WP_THREADS + t;
  const int sh = w < nw ? ldg(S.w_shell + w) : -1;
  double epol = 0.0, th = 0.0;
  if (sh >= 0) {
    const double* st = S.sh_state;
    const int n = S.sh_n[sh];
    const int il = ldg(S.w_rank + w) + 1;
    th = ldg(S.w_theta + w);
    const double tc = st[sh];
    double arg = 1.0 + (1.0 - 2.0 * (double)il) / (double)n;
    double th0 = acos(arg);
    th0 = th0 - 3.0 * sin(th0) * S.sh_cstb[sh] / 2.0;
    if (th0 < 0.0) th0 = 0.0;
    if (th0 > pi) th0 = pi;
    const double dth = th - th0 + tc;
    epol = 0.5 * S.fkwpol * (dth * dth);
    const double dv = S.fkwpol * dth;
    const int i = S.nat_solute + S.sa * w;
    double rmu[3], rcu[3], rm, rc;
    wat_vectors(S, i, rmu, rm, rcu, rc);
    double scp = rmu[0] * rcu[0] + rmu[1] * rcu[1] + rmu[2] * rcu[2];
    scp = scp > 1.0 ? 1.0 : (scp < -1.0 ? -1.0 : scp);
    double f0 = sin(acos(scp));
    if (fabs(f0) < 1.0e-10) f0 = 1.0e-10;
    f0 = -1.0 / f0;
    f0 = dv * f0;
    const auto d = gpm(S.d);
    for (int c = 0; c < 3; c++) {
      const double av = rcu[c] - rmu[c] * scp;
      const double f1 = (av * (-2.0)) / rm, f3 = av / rm, f2 = (rmu[c] - rcu[c] * scp) / rc;
      d[3 * i + c] += (f1 + f2) * f0;
      for (int k = 1; k < S.sa; k++) d[3 * (i + k) + c] += f3 * f0;
    }
  }
  // block sums: [0] energy, [1 + is] theta over the block's members of shell is
  __shared__ double red[WP_THREADS / QNB_WAVE][1 + QNB_MAXSH];
  double v = wave_sum(epol);
  if (lane == 0) red[wid][0] = v;
  for (int is = 0; is < nsh; is++) {
    const double s2 = wave_sum(sh == is ? th : 0.0);
    if (lane == 0) red[wid][1 + is] = s2;
  }
  __syncthreads();
  if (t == 0) {
    double* part = S.wp_part + (size_t)blockIdx.x * (1 + QNB_MAXSH);
    for (int c = 0; c <= nsh; c++) {
      double s2 = 0.0;
      for (int k = 0; k < WP_THREADS / QNB_WAVE; k++) s2 += red[k][c];
      part[c] = s2;
    }
  }
}

// Phase 4, one thread per system: adds the block sums in block order,
// accumulates avtheta and avn_insh, writes the energy slot and clears the
// shell counts for the next step's phase 1.
__global__ __launch_bounds__(64) void k_watpol_finish(const DevSys* __restrict__ systems) {
  const DevSys& S = systems[blockIdx.x];
  if (threadIdx.x != 0) return;
  double* ep = S.epart + (size_t)(S.ncgp + 1) * QNB_NE;
  for (int k = 0; k < QNB_NE; k++) ep[k] = 0.0;
  if (!has_pol(S)) return;
  const int nsh = S.nshell, nblk = (S.nwat + WP_THREADS - 1) / WP_THREADS;
  double sum[1 + QNB_MAXSH];
  for (int k = 0; k <= nsh; k++) sum[k] = 0.0;
  for (int blk = 0; blk < nblk; blk++) {
    const double* pb = S.wp_part + (size_t)blk * (1 + QNB_MAXSH);
    for (int k = 0; k <= nsh; k++) sum[k] += pb[k];
  }
  double* st = S.sh_state;
  for (int is = 0; is < nsh; is++) {
    const int n = S.sh_n[is];
    S.sh_n[is] = 0;
    if (n == 0) continue;
    st[QNB_MAXSH + is] += sum[1 + is] / (double)n;
    st[2 * QNB_MAXSH + is] += (double)n;
  }
  ep[E_POL] = sum[0];
}

// ---------------------------------------------------------------------------
// Fixed-order reduction of the per-wave energy partials: wave c of the block
// sums slots [c*n/16, (c+1)*n/16) with lane k on energy k (a wave reads one
// 57-double slot per iteration), then thread k adds the 16 chunk sums in
// order.  Runs on the sampled steps only.
#define QNB_RED_CHUNKS 16
__global__ __launch_bounds__(64 * QNB_RED_CHUNKS) void k_reduce_energies(const DevSys* __restrict__ systems,
                                                                        int with_qq, int with_pol) {
  const DevSys& S = systems[blockIdx.x];
  if (!want_energies(S)) return;
  __shared__ double part[QNB_RED_CHUNKS][QNB_WAVE];
  const int k = threadIdx.x % QNB_WAVE, c = threadIdx.x / QNB_WAVE;
  const int n = S.ncgp, a0 = (int)((long)n * c / QNB_RED_CHUNKS), a1 = (int)((long)n * (c + 1) / QNB_RED_CHUNKS);
  double s = 0.0;
  if (k < QNB_NE)
    for (int a = a0; a < a1; a++) s += S.epart[(size_t)a * QNB_NE + k];
  part[c][k] = s;
  // per-atom LRF and radial-restraint energies: thread t sums a contiguous
  // atom range, then the ranges are added in order
  __shared__ double apart[64 * QNB_RED_CHUNKS][2];
  {
    const int t = threadIdx.x, nt = blockDim.x;
    const int i0 = (int)((long)S.natom * t / nt), i1 = (int)((long)S.natom * (t + 1) / nt);
    double l = 0.0, r = 0.0;
    for (int i = i0; i < i1; i++) {
      l += S.eatom[2 * i];
      r += S.eatom[2 * i + 1];
    }
    apart[t][0] = l;
    apart[t][1] = r;
  }
  __syncthreads();
  if (c == 0 && k < QNB_NE) {
    double t = 0.0;
    for (int m = 0; m < QNB_RED_CHUNKS; m++) t += part[m][k];
    if (k == E_LRF || k == E_RAD) {
      const int col = k == E_LRF ? 0 : 1;
      for (int m = 0; m < (int)blockDim.x; m++) t += apart[m][col];
    }
    if (with_qq && S.nqat > 0) t += S.epart[(size_t)S.ncgp * QNB_NE + k];
    if (with_pol && S.nwat > 0 && S.wpol && S.nshell > 0) t += S.epart[(size_t)(S.ncgp + 1) * QNB_NE + k];
    S.ered[k] = t;
  }
}

// explicit instantiations used by the host code
template __global__ void k_lists_lrf<double, false>(const DevSys*);
template __global__ void k_lists_lrf<float, false>(const DevSys*);
template __global__ void k_lists_lrf<double, true>(const DevSys*);
template __global__ void k_lists_lrf<float, true>(const DevSys*);
#define QNB_INST(R, NS)                                                            \
  template __global__ void k_qforces<R, NS>(const DevSys*, uint32_t, int, double*, int);                       \
  template __global__ void k_sforces<R, NS, true, WF_CAP_SMALL>(const DevSys*, uint32_t, int, int, int); \
  template __global__ void k_sforces<R, NS, false, WF_CAP_SMALL>(const DevSys*, uint32_t, int, int, int); \
  template __global__ void k_sforces<R, NS, true, FC_CAP>(const DevSys*, uint32_t, int, int, int);       \
  template __global__ void k_sforces<R, NS, false, FC_CAP>(const DevSys*, uint32_t, int, int, int);      \
  template __global__ void k_wforces<R, NS, true, WF_CAP_SMALL>(const DevSys*, uint32_t, int, int, int); \
  template __global__ void k_wforces<R, NS, false, WF_CAP_SMALL>(const DevSys*, uint32_t, int, int, int); \
  template __global__ void k_wforces<R, NS, true, FC_CAP>(const DevSys*, uint32_t, int, int, int);       \
  template __global__ void k_wforces<R, NS, false, FC_CAP>(const DevSys*, uint32_t, int, int, int);       \
  template __global__ void k_gforces<R, NS>(const DevSys*, uint32_t, int, int, int);
QNB_INST(double, 1)
QNB_INST(double, 2)
QNB_INST(double, 4)
QNB_INST(double, 8)
QNB_INST(float, 1)
QNB_INST(float, 2)
QNB_INST(float, 4)
QNB_INST(float, 8)
#define QNB_INST_QQ(NS)                                           \
  template __global__ void k_qq<double, NS>(const DevSys*, uint32_t); \
  template __global__ void k_qq<float, NS>(const DevSys*, uint32_t);
QNB_INST_QQ(1)
QNB_INST_QQ(2)
QNB_INST_QQ(4)
QNB_INST_QQ(8)

}  // namespace qnb
