// gr_adjoint.hip -- reverse sweep of the DEM step (FP64, physics semantics,
// spheres and ellipsoids, Rectangle walls, a fixed contact set per step).
//
// One backward step t, given the adjoint of the state at t+1 (13 planes:
// pos 3, quat 4, vel 3, omega 3) and the state at t (the forward step t has
// just been re-run, so the entries, their contact flags and the total force
// and torque of step t are in place):
//   k_adj_particle  integration adjoint -> the particle's own partial state
//                   adjoint and the adjoint of its total force and torque;
//                   wall contacts (per particle, as in the forward)
//   k_adj_entry     per contact entry (i, j): geometry + Hooke adjoint ->
//                   partials for i and for j and the entry's parameter
//                   adjoints (no atomics)
//   k_adj_gather    per particle: own partial + its entries' i-partials +
//                   the j-partials of the mirror entries (j, i) in j's list,
//                   summed in a fixed order
//   k_adj_params    one block per contact-table entry: the parameter
//                   adjoints of the step, reduced in a fixed order
// Every sum has a fixed order: the gradients are reproducible bit for bit.
#include "gr_adjoint.hpp"
#include "gr_internal.hpp"

namespace gr {

using namespace adj;

namespace {

// 13 planes of n doubles: pos (0..2), quat (3..6), vel (7..9), omega (10..12)
struct P13 {
  double* p;
  uint64_t n;
  __device__ V3<double> v3(int plane, uint64_t i) const { return {p[plane * n + i], p[(plane + 1) * n + i], p[(plane + 2) * n + i]}; }
  __device__ Qt<double> q(uint64_t i) const { return {p[3 * n + i], p[4 * n + i], p[5 * n + i], p[6 * n + i]}; }
  __device__ void put3(int plane, uint64_t i, V3<double> a) const {
    p[plane * n + i] = a.x;
    p[(plane + 1) * n + i] = a.y;
    p[(plane + 2) * n + i] = a.z;
  }
  __device__ void putq(uint64_t i, Qt<double> a) const {
    p[3 * n + i] = a.x;
    p[4 * n + i] = a.y;
    p[5 * n + i] = a.z;
    p[6 * n + i] = a.w;
  }
};

// per-entry partial state adjoint: 13 doubles (x, q, v, w)
struct Part {
  V3<double> x;
  Qt<double> q;
  V3<double> v, w;
};
__device__ void put_part(double* base, uint64_t e, const Part& a) {
  double* o = base + 13 * e;
  o[0] = a.x.x, o[1] = a.x.y, o[2] = a.x.z;
  o[3] = a.q.x, o[4] = a.q.y, o[5] = a.q.z, o[6] = a.q.w;
  o[7] = a.v.x, o[8] = a.v.y, o[9] = a.v.z;
  o[10] = a.w.x, o[11] = a.w.y, o[12] = a.w.z;
}
__device__ void add_part(double* acc, const double* base, uint64_t e) {
  const double* o = base + 13 * e;
#pragma unroll
  for (int k = 0; k < 13; ++k) acc[k] += o[k];
}

__device__ Tr<double> pose_of(V3<double> x, Qt<double> q) {
  Tr<double> t;
  t.R = rot(q);
  t.o = x;
  return t;
}

// Hooke parameters of the table entry
__device__ HookeIn hooke_in(const HookeP<double>& h) { return {h.kn, h.muen, h.etat, h.muc}; }

}  // namespace

// ---- integration and walls -------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_adj_particle(DevScene d, AdjBufs a) {
  const uint64_t n = d.n;
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const P13 st{a.state, n}, nb{a.abar, n}, own{a.own, n};
  const V3<double> x = st.v3(0, i), v = st.v3(7, i), w = st.v3(10, i);
  const Qt<double> q = st.q(i);
  const V3<double> Mt = load3(d.last_torque, n, i);
  const uint32_t si = d.shape[i];
  const ShapeDyn dy = d.dyn[si];
  StateBar s1, s0;
  s1.x = nb.v3(0, i);
  s1.q = nb.q(i);
  s1.v = nb.v3(7, i);
  s1.w = nb.v3(10, i);
  V3<double> Fb, Mb;
  integrate_vjp(Body{dy.mass, dy.inertia, dy.inv}, d.dt, x, q, v, w, Mt, s1, s0, Fb, Mb);
  // walls (the forward's particle x obstacle loop, ComponentManagerCPU.cpp:301-351);
  // every slot of the particle is cleared, its sim may hold fewer walls
  for (uint32_t k = 0; k < a.max_obs; ++k) a.wall_tidx[i * a.max_obs + k] = 0xffffffffu;
  const SimGrid& g = d.grids[d.sim[i]];
  if (g.obs_end > g.obs_begin) {
    const ShapeNP<double> A = d.np64[si];
    const Tr<double> a2w = pose_of(x, q);
    PoseBar pb = posebar_zero();
    for (uint32_t o = g.obs_begin; o < g.obs_end; ++o) {
      const uint64_t slot = i * a.max_obs + (o - g.obs_begin);
      const ShapeNP<double> B = d.np64[d.obs_shape[o]];
      const double* t = d.obs_tr + 12 * (size_t)o;
      Tr<double> b2w;
#pragma unroll
      for (int k = 0; k < 9; ++k) b2w.R.m[k] = t[k];
      b2w.o = {t[9], t[10], t[11]};
      const CInfo<double> ci = cp_rectangle(sup_ctx(A, false, d.verts64, d.cverts64), a2w, b2w);
      if (!(ci.overlap < 0.0)) continue;
      const uint32_t tix = g.table + material_hash((uint32_t)A.material, (uint32_t)B.material);
      const HookeP<double> h = d.hooke64[tix];
      HookeBar hb = hookebar_zero();
      hooke_vjp(hooke_in(h), ci.cv, ci.cp, v, w, V3<double>{0, 0, 0}, V3<double>{0, 0, 0}, x, ci.cp, A.mass,
                B.mass, Fb, Mb, hb);
      V3<double> rN = {b2w.R.m[2], b2w.R.m[5], b2w.R.m[8]};
      rN = copysign(1.0, dot(rN, x - b2w.o)) * rN;
      wall_geo_vjp(A.kind, A.dims, Pose{x, a2w.R}, rN, hb.cv, hb.cp, pb);
      s0.x = s0.x + hb.xA;
      s0.v = s0.v + hb.vA;
      s0.w = s0.w + hb.wA;
      a.wall_tidx[slot] = tix;
      double* tb = a.wall_theta + 4 * slot;
      tb[0] = hb.kn;
      tb[1] = hb.muen * dmuen_den(h.en);
      tb[2] = hb.etat;
      tb[3] = hb.muc;
    }
    s0.x = s0.x + pb.x;
    s0.q = qadd(s0.q, rot_vjp(q, pb.R));
  }
  own.put3(0, i, s0.x);
  own.putq(i, s0.q);
  own.put3(7, i, s0.v);
  own.put3(10, i, s0.w);
  store3(a.Fb, n, i, Fb);
  store3(a.Mb, n, i, Mb);
}

// ---- contacts --------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_adj_entry(DevScene d, AdjBufs a) {
  const uint64_t n = d.n;
  const uint64_t E = d.entry_off[n];
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const P13 st{a.state, n};
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += stride) {
    Part pi = {{0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, pj = pi;
    a.entry_tidx[e] = 0xffffffffu;
    if (d.pair_flag[e] & kFlagContact) {
      const uint32_t i = d.sidx[d.pair_i[e]], j = d.sidx[d.pair_j[e]];
      const ShapeNP<double> A = d.np64[d.shape[i]], B = d.np64[d.shape[j]];
      const SimGrid& g = d.grids[d.sim[i]];
      const V3<double> xi = st.v3(0, i), xj0 = st.v3(0, j);
      const V3<double> dd = xj0 - xi;
      const V3<double> dm = g.periodic ? min_image(g, dd) : dd;
      const V3<double> xj = same(dm, dd) ? xj0 : xi + dm;
      const Qt<double> qi = st.q(i), qj = st.q(j);
      const Tr<double> a2w = pose_of(xi, qi), b2w = pose_of(xj, qj);
      const GjkArgs args = {d.flags, d.gjk_rel};
      // forward geometry, as closestPointsRigidBodies with physics semantics
      CInfo<double> ci = no_contact<double>();
      int branch = 0;  // 1 spheres, 2 deep cores, 3 separated cores
      V3<double> nrm = {0, 0, 0};
      double dist = 0.0;
      const double ctSum = A.crust + B.crust;
      if (A.kind == GRAINS_SPHERE && B.kind == GRAINS_SPHERE) {
        ci = cp_spheres(A.rc, B.rc, a2w, b2w);
        branch = 1;
      } else {
        SupCtx<double> ca = sup_ctx(A, true, d.verts64, d.cverts64);
        SupCtx<double> cb = sup_ctx(B, true, d.verts64, d.cverts64);
        const GjkOut<double> go = gjk_sv(ca, cb, a2w, b2w, args, A.rc + B.rc, ctSum);
        if (!go.separated) {
          double distance = go.dist - ctSum;
          if (go.deep) distance = -ctSum;
          if (!(distance > 0.0)) {
            const V3<double> ptA = apply(a2w, go.pa), ptB = apply(b2w, go.pb);
            ci.cp = (ptA + ptB) / 2.0;
            V3<double> cv = ptA - ptB;
            if (go.deep) {
              cv = a2w.o - b2w.o;
              ci.cp = (a2w.o + b2w.o) / 2.0;
              branch = 2;
            } else {
              branch = 3;
              dist = go.dist;
              nrm = (ptB - ptA) / norm(ptB - ptA);
            }
            cv = cv / norm(cv);
            cv = round_tol(cv, Eps<double>::MID);
            ci.cv = (-distance) * cv;
            ci.overlap = distance;
          }
        }
      }
      if (ci.overlap < 0.0) {
        const uint32_t tix = g.table + material_hash((uint32_t)A.material, (uint32_t)B.material);
        const HookeP<double> h = d.hooke64[tix];
        const V3<double> Fb = load3(a.Fb, n, i), Mb = load3(a.Mb, n, i);
        HookeBar hb = hookebar_zero();
        hooke_vjp(hooke_in(h), ci.cv, ci.cp, st.v3(7, i), st.v3(10, i), st.v3(7, j), st.v3(10, j), xi, xj, A.mass,
                  B.mass, Fb, Mb, hb);
        V3<double> xAb = hb.xA, xBb = hb.xB;
        PoseBar Ab = posebar_zero(), Bb = posebar_zero();
        if (branch == 1) {
          sphere_geo_vjp(xi, xj, A.rc, B.rc, hb.cv, hb.cp, xAb, xBb);
        } else if (branch == 2) {
          deep_geo_vjp(xi, xj, ctSum, hb.cv, hb.cp, xAb, xBb);
        } else {
          gjk_geo_vjp(A.kind, A.core, Pose{xi, a2w.R}, B.kind, B.core, Pose{xj, b2w.R}, nrm, dist, ctSum, hb.cv,
                      hb.cp, Ab, Bb);
        }
        pi.x = xAb + Ab.x;
        pi.q = rot_vjp(qi, Ab.R);
        pi.v = hb.vA;
        pi.w = hb.wA;
        pj.x = xBb + Bb.x;
        pj.q = rot_vjp(qj, Bb.R);
        pj.v = hb.vB;
        pj.w = hb.wB;
        a.entry_tidx[e] = tix;
        double* tb = a.entry_theta + 4 * e;
        tb[0] = hb.kn;
        tb[1] = hb.muen * dmuen_den(h.en);
        tb[2] = hb.etat;
        tb[3] = hb.muc;
      }
    }
    put_part(a.part_i, e, pi);
    put_part(a.part_j, e, pj);
  }
}

// ---- gather ----------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_adj_gather(DevScene d, AdjBufs a) {
  const uint64_t n = d.n;
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const P13 own{a.own, n}, out{a.abar, n};
  double acc[13];
#pragma unroll
  for (int k = 0; k < 13; ++k) acc[k] = a.own[k * n + i];
  const uint32_t s = d.rank[i];
  for (uint32_t e = d.entry_off[s]; e < d.entry_off[s + 1]; ++e) {
    add_part(acc, a.part_i, e);
    // the mirror entry (pair_j[e], s) in its partner's list
    const uint32_t sj = d.pair_j[e];
    for (uint32_t f = d.entry_off[sj]; f < d.entry_off[sj + 1]; ++f)
      if (d.pair_j[f] == s) {
        add_part(acc, a.part_j, f);
        break;
      }
  }
#pragma unroll
  for (int k = 0; k < 13; ++k) a.abar[k * n + i] = acc[k];
  (void)own;
  (void)out;
}

// ---- parameters --------------------------------------------------------------------
// Block t sums the step's adjoints of table entry t (entries, then walls) in a
// fixed order and adds them to dtheta[4 t ...].
__global__ void __launch_bounds__(kBlock) k_adj_params(DevScene d, AdjBufs a) {
  __shared__ double red[kBlock];
  const uint32_t t = blockIdx.x;
  const uint64_t E = d.entry_off[d.n];
  const uint64_t W = d.n * a.max_obs;
  for (int k = 0; k < 4; ++k) {
    double s = 0.0;
    for (uint64_t e = threadIdx.x; e < E; e += blockDim.x)
      if (a.entry_tidx[e] == t) s += a.entry_theta[4 * e + k];
    for (uint64_t w = threadIdx.x; w < W; w += blockDim.x)
      if (a.wall_tidx[w] == t) s += a.wall_theta[4 * w + k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) a.dtheta[4 * t + k] += red[0];
    __syncthreads();
  }
}

void launch_adjoint_step(const DevScene& d, const AdjBufs& a, uint32_t n_tables, hipStream_t s) {
  const unsigned gp = (unsigned)((d.n + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_adj_particle, dim3(gp), dim3(kBlock), 0, s, d, a);
  hipLaunchKernelGGL(k_adj_entry, dim3(d.narrow_blocks), dim3(kBlock), 0, s, d, a);
  hipLaunchKernelGGL(k_adj_gather, dim3(gp), dim3(kBlock), 0, s, d, a);
  hipLaunchKernelGGL(k_adj_params, dim3(n_tables), dim3(kBlock), 0, s, d, a);
}

}  // namespace gr
