// FP64 integrators over packed molecules, in TensorMol's MD units
// (A, fs, kg/mol; forces J/mol/A).
//
// Velocity Verlet (TensorMol/Simulations/SimpleMD.py:14-38):
//   x <- x + v dt + a dt^2 / 2 ;  F(x) ;  a' = 1e-10 F / m ;  v <- v + (a + a') dt / 2
// Nose-Hoover, single thermostat variable per molecule (SimpleMD.py:104-129):
//   x <- x + v dt + (a - eta v) dt^2 / 2 ; vh = v + (a - eta v) dt / 2 ; F(x) ; a' = 1e-10 F/m
//   eta <- eta + dt/(2Q) (KE(v) - (3N+1)/2 kT) + dt/(2Q) (KE(vh) - (3N+1)/2 kT)
//   v <- (vh + a' dt / 2) / (1 + eta dt / 2)
// FIRE (Bitzek et al., PRL 97, 170201 (2006)); the reference has no FIRE.
#include "tensormol_amd.h"
#include "tm_common.h"

namespace {

__global__ void vv_drift_kernel(int N, double dt, const double* __restrict__ v,
                                const double* __restrict__ a, double* __restrict__ x) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * N) return;
  x[t] = x[t] + v[t] * dt + 0.5 * a[t] * dt * dt;
}

// a' = 1e-10 F / m ; v += (a + a') dt / 2 ; a = a' ; ke_atom = m |v|^2 / 2
__global__ void vv_kick_kernel(int N, double dt, const double* __restrict__ f,
                               const double* __restrict__ m, double* __restrict__ v,
                               double* __restrict__ a, double* __restrict__ ke) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double mi = m[i], s = 0.0;
  for (int d = 0; d < 3; ++d) {
    double an = 1e-10 * f[3 * i + d] / mi;
    double vn = v[3 * i + d] + 0.5 * (a[3 * i + d] + an) * dt;
    v[3 * i + d] = vn;
    a[3 * i + d] = an;
    s += vn * vn;
  }
  if (ke) ke[i] = 0.5 * mi * s;
}

// vv_kick of one step and vv_drift of the next in one pass over the atoms:
// the same operations on the same values as the two kernels in turn
__global__ void vv_kick_drift_kernel(int N, double dt, const double* __restrict__ f,
                                     const double* __restrict__ m, double* __restrict__ v,
                                     double* __restrict__ a, double* __restrict__ ke,
                                     double* __restrict__ x) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double mi = m[i], s = 0.0;
  for (int d = 0; d < 3; ++d) {
    double an = 1e-10 * f[3 * i + d] / mi;
    double vn = v[3 * i + d] + 0.5 * (a[3 * i + d] + an) * dt;
    v[3 * i + d] = vn;
    a[3 * i + d] = an;
    s += vn * vn;
    x[3 * i + d] = x[3 * i + d] + vn * dt + 0.5 * an * dt * dt;
  }
  if (ke) ke[i] = 0.5 * mi * s;
}

// Nose-Hoover first half: x, vh from (x, v, a, eta[mol]); ke2[i] = (m|v|^2/2, m|vh|^2/2)
__global__ void nh_drift_kernel(int N, double dt, const int* __restrict__ mol,
                                const double* __restrict__ eta, const double* __restrict__ m,
                                double* __restrict__ x, const double* __restrict__ v,
                                const double* __restrict__ a, double* __restrict__ vh,
                                double* __restrict__ ke2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double e = eta[mol[i]], s0 = 0.0, s1 = 0.0;
  for (int d = 0; d < 3; ++d) {
    int t = 3 * i + d;
    double vv = v[t], acc = a[t] - e * vv;
    x[t] = x[t] + vv * dt + 0.5 * acc * dt * dt;
    double h = vv + 0.5 * acc * dt;
    vh[t] = h;
    s0 += vv * vv;
    s1 += h * h;
  }
  ke2[2 * i] = 0.5 * m[i] * s0;
  ke2[2 * i + 1] = 0.5 * m[i] * s1;
}

// per molecule: eta update from the two kinetic energies (ke2sum[B,2])
__global__ void nh_eta_kernel(int B, const int* __restrict__ mol_start, double dt,
                              const double* __restrict__ kT, const double* __restrict__ Q,
                              const double* __restrict__ ke2sum, double* __restrict__ eta) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  double n = (double)(mol_start[b + 1] - mol_start[b]);
  double target = ((3.0 * n + 1.0) / 2.0) * kT[b];
  double q = Q[b];
  double e = eta[b] + (dt / (2.0 * q)) * (ke2sum[2 * b] - target);
  e = e + (dt / (2.0 * q)) * (ke2sum[2 * b + 1] - target);
  eta[b] = e;
}

__global__ void nh_kick_kernel(int N, double dt, const int* __restrict__ mol,
                               const double* __restrict__ eta, const double* __restrict__ f,
                               const double* __restrict__ m, const double* __restrict__ vh,
                               double* __restrict__ v, double* __restrict__ a,
                               double* __restrict__ ke) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double e = eta[mol[i]], mi = m[i], s = 0.0;
  for (int d = 0; d < 3; ++d) {
    int t = 3 * i + d;
    double an = 1e-10 * f[t] / mi;
    double vn = (vh[t] + 0.5 * dt * an) / (1.0 + 0.5 * dt * e);
    v[t] = vn;
    a[t] = an;
    s += vn * vn;
  }
  if (ke) ke[i] = 0.5 * mi * s;
}

// FIRE, per atom partial sums for the per-molecule power P = F.v, |F|^2, |v|^2
__global__ void fire_dots_kernel(int N, const double* __restrict__ f, const double* __restrict__ v,
                                 double* __restrict__ dots) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double p = 0, ff = 0, vv = 0;
  for (int d = 0; d < 3; ++d) {
    double fd = f[3 * i + d], vd = v[3 * i + d];
    p += fd * vd; ff += fd * fd; vv += vd * vd;
  }
  dots[3 * i] = p; dots[3 * i + 1] = ff; dots[3 * i + 2] = vv;
}

// FIRE state update per molecule: state[b] = (dt, alpha, n_pos, active)
__global__ void fire_state_kernel(int B, const double* __restrict__ dots_sum,
                                  double* __restrict__ state, double* __restrict__ mix,
                                  double dtmax, int nmin, double finc, double fdec,
                                  double alpha0, double falpha) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  double* st = state + 4 * b;
  double P = dots_sum[3 * b], ff = dots_sum[3 * b + 1], vv = dots_sum[3 * b + 2];
  double dt = st[0], al = st[1], npos = st[2];
  double c_v = 1.0 - al, c_f = 0.0, zero_v = 0.0;
  if (P > 0.0) {
    c_f = ff > 0.0 ? al * sqrt(vv / ff) : 0.0;
    npos += 1.0;
    if (npos > nmin) {
      dt = fmin(dt * finc, dtmax);
      al *= falpha;
    }
  } else {
    dt *= fdec;
    al = alpha0;
    npos = 0.0;
    zero_v = 1.0;
  }
  if (st[3] == 0.0) {   // converged molecules do not move
    dt = 0.0;
    zero_v = 1.0;
  }
  st[0] = dt; st[1] = al; st[2] = npos;
  mix[3 * b] = c_v; mix[3 * b + 1] = c_f; mix[3 * b + 2] = zero_v;
}

// v <- (1-alpha) v + alpha |v| F/|F| (or 0), then semi-implicit Euler:
// v += dt F / m_eff ; x += dt v  (unit mass scaled to A and fs as in MD)
__global__ void fire_move_kernel(int N, const int* __restrict__ mol, int mol0,
                                 const double* __restrict__ mix, const double* __restrict__ state,
                                 const double* __restrict__ f, const double* __restrict__ m,
                                 double* __restrict__ v, double* __restrict__ x, double maxstep) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int b = mol[i] - mol0;
  double c_v = mix[3 * b], c_f = mix[3 * b + 1], zero_v = mix[3 * b + 2];
  double dt = state[4 * b];
  double dx[3], n2 = 0.0;
  for (int d = 0; d < 3; ++d) {
    int t = 3 * i + d;
    double vn = zero_v != 0.0 ? 0.0 : c_v * v[t] + c_f * f[t];
    vn += dt * 1e-10 * f[t] / m[i];
    v[t] = vn;
    dx[d] = dt * vn;
    n2 += dx[d] * dx[d];
  }
  double sc = 1.0;
  if (n2 > maxstep * maxstep) sc = maxstep / sqrt(n2);
  for (int d = 0; d < 3; ++d) x[3 * i + d] += sc * dx[d];
}

// Result staging of a serving step: several device arrays (whole 4-byte
// words) concatenated into one buffer, one array per blockIdx.y.
constexpr int PACK_PARTS = 16;

struct PackBatch {
  tm_pack_part p[PACK_PARTS];
};

__global__ void pack_copy_kernel(const PackBatch pb, unsigned* __restrict__ dst) {
  const tm_pack_part& q = pb.p[blockIdx.y];
  const unsigned* src = static_cast<const unsigned*>(q.src);
  unsigned* out = dst + q.offset / 4;
  const long n = q.bytes / 4;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long)gridDim.x * blockDim.x)
    out[t] = src[t];
}

}  // namespace

extern "C" {

int tm_vv_drift(int N, double dt, const double* v, const double* a, double* x, void* stream) {
  if (N == 0) return TM_OK;
  vv_drift_kernel<<<tm_div_up(3l * N, 256), 256, 0, (hipStream_t)stream>>>(N, dt, v, a, x);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

int tm_vv_kick(int N, double dt, const double* f, const double* m, double* v, double* a,
               double* ke, void* stream) {
  if (N == 0) return TM_OK;
  vv_kick_kernel<<<tm_div_up(N, 256), 256, 0, (hipStream_t)stream>>>(N, dt, f, m, v, a, ke);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

int tm_vv_kick_drift(int N, double dt, const double* f, const double* m, double* v, double* a,
                     double* ke, double* x, void* stream) {
  if (N == 0) return TM_OK;
  vv_kick_drift_kernel<<<tm_div_up(N, 256), 256, 0, (hipStream_t)stream>>>(N, dt, f, m, v, a, ke,
                                                                           x);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

int tm_nh_drift(int N, double dt, const int* mol, const double* eta, const double* m, double* x,
                const double* v, const double* a, double* vh, double* ke2, void* stream) {
  if (N == 0) return TM_OK;
  nh_drift_kernel<<<tm_div_up(N, 256), 256, 0, (hipStream_t)stream>>>(N, dt, mol, eta, m, x, v, a,
                                                                      vh, ke2);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

int tm_nh_eta(int B, const int* mol_start, double dt, const double* kT, const double* Q,
              const double* ke2sum, double* eta, void* stream) {
  if (B == 0) return TM_OK;
  nh_eta_kernel<<<tm_div_up(B, 128), 128, 0, (hipStream_t)stream>>>(B, mol_start, dt, kT, Q,
                                                                    ke2sum, eta);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

int tm_nh_kick(int N, double dt, const int* mol, const double* eta, const double* f,
               const double* m, const double* vh, double* v, double* a, double* ke, void* stream) {
  if (N == 0) return TM_OK;
  nh_kick_kernel<<<tm_div_up(N, 256), 256, 0, (hipStream_t)stream>>>(N, dt, mol, eta, f, m, vh, v,
                                                                     a, ke);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

int tm_fire_dots(int N, const double* f, const double* v, double* dots, void* stream) {
  if (N == 0) return TM_OK;
  fire_dots_kernel<<<tm_div_up(N, 256), 256, 0, (hipStream_t)stream>>>(N, f, v, dots);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

// Atoms [0, N) of the pointers belong to molecules mol0 .. mol0+B-1 (mol[] holds
// global molecule ids); dots_sum, state and mix are indexed by mol - mol0.
int tm_fire_update(int N, int B, const int* mol, int mol0, const double* dots_sum, double* state,
                   double* mix, const double* f, const double* m, double* v, double* x,
                   double dtmax, int nmin, double finc, double fdec, double alpha0,
                   double falpha, double maxstep, void* stream) {
  if (N == 0 || B == 0) return TM_OK;
  hipStream_t s = (hipStream_t)stream;
  fire_state_kernel<<<tm_div_up(B, 128), 128, 0, s>>>(B, dots_sum, state, mix, dtmax, nmin, finc,
                                                      fdec, alpha0, falpha);
  TM_LAUNCH_CHECK();
  fire_move_kernel<<<tm_div_up(N, 256), 256, 0, s>>>(N, mol, mol0, mix, state, f, m, v, x, maxstep);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

int tm_pack_copy(int nparts, const tm_pack_part* parts, void* dst, void* stream) {
  TM_REQUIRE(nparts >= 0 && (nparts == 0 || (parts != nullptr && dst != nullptr)),
             "tm_pack_copy: bad part list");
  for (int i = 0; i < nparts; ++i)
    TM_REQUIRE(parts[i].bytes >= 0 && parts[i].offset >= 0 && parts[i].bytes % 4 == 0 &&
                   parts[i].offset % 4 == 0 && (parts[i].bytes == 0 || parts[i].src != nullptr),
               "tm_pack_copy: part %d: sizes and offsets must be whole 4-byte words", i);
  for (int g0 = 0; g0 < nparts;) {
    PackBatch pb;
    int n = 0;
    long most = 0;
    for (; g0 < nparts && n < PACK_PARTS; ++g0) {
      if (parts[g0].bytes == 0) continue;
      pb.p[n++] = parts[g0];
      most = most > parts[g0].bytes / 4 ? most : parts[g0].bytes / 4;
    }
    if (n == 0) continue;
    const int blocks = (int)(most > 256 * 255 ? 256 : tm_div_up(most, 256));
    pack_copy_kernel<<<dim3(blocks, n), 256, 0, (hipStream_t)stream>>>(
        pb, static_cast<unsigned*>(dst));
    TM_LAUNCH_CHECK();
  }
  return TM_OK;
}

}  // extern "C"
