/*
 * pxd_oracle.c — FP64 CPU restatement of the Protenix-Dock docking hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see pxd_oracle.h).  The algorithms below follow
 * the reference's semantics term by term; each function names the reference
 * file:line it restates.  Arithmetic is IEEE double like the reference's
 * default build (param_t = double, engine/CMakeLists.txt:17-20).
 */
#include "pxd_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "pxdock_gbsa_table.h"
#include "pxdock_vdw_table.h"

/* system.h:83 */
const double or_coulomb_factor = 332.0637141491396;
/* math.h:128 */
#define EPS 1e-20
/* math.h:226-227 */
#define PI 3.141592653589793

/* constant.h:29-130 */
#define SIGMA_ROW(i, tag, s, e) s,
#define EPSILON_ROW(i, tag, s, e) e,
const double or_gaff2_sigma[PXD_NUM_VDW_TYPES] = {PXD_VDW_TABLE(SIGMA_ROW)};
const double or_gaff2_epsilon[PXD_NUM_VDW_TYPES] = {PXD_VDW_TABLE(EPSILON_ROW)};

/* ======================================================================== */
/* small vector helpers (math.h:54-200)                                     */

static inline void sub3(const double* a, const double* b, double* c) {
  c[0] = a[0] - b[0];
  c[1] = a[1] - b[1];
  c[2] = a[2] - b[2];
}
static inline double dot3(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline void cross3(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
static inline double sqsum3(const double* a) { return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]; }
static inline double norm3(const double* a) { return sqrt(sqsum3(a)); }
static inline void scale3(double* a, double f) {
  a[0] *= f;
  a[1] *= f;
  a[2] *= f;
}
static inline void axpy3(double* y, double f, const double* x) {
  y[0] += f * x[0];
  y[1] += f * x[1];
  y[2] += f * x[2];
}
static inline void mv3(const double* m, const double* v, double* o) {
  o[0] = m[0] * v[0] + m[1] * v[1] + m[2] * v[2];
  o[1] = m[3] * v[0] + m[4] * v[1] + m[5] * v[2];
  o[2] = m[6] * v[0] + m[7] * v[1] + m[8] * v[2];
}
/* math.h:176-178 */
static inline double safe_div(double a, double b) { return a / (b + EPS); }
/* math.h:180-182 (clamp) */
static inline double clampd(double v, double lo, double hi) {
  double t = v > lo ? v : lo;
  return t < hi ? t : hi;
}

/* math.h:38-49 */
void or_matmul3(const double* r, const double* s, double* out) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      out[3 * i + j] = r[3 * i] * s[j] + r[3 * i + 1] * s[3 + j] + r[3 * i + 2] * s[6 + j];
}

/* math.cpp:22-42: quaternion (a, b, c, d) to rotation matrix */
void or_rotmat_quat(const double q[4], double m[9]) {
  double aa = q[0] * q[0], ab = q[0] * q[1], ac = q[0] * q[2], ad = q[0] * q[3];
  double bb = q[1] * q[1], bc = q[1] * q[2], bd = q[1] * q[3];
  double cc = q[2] * q[2], cd = q[2] * q[3], dd = q[3] * q[3];
  m[0] = aa + bb - cc - dd;
  m[1] = 2 * (bc - ad);
  m[2] = 2 * (ac + bd);
  m[3] = 2 * (ad + bc);
  m[4] = aa - bb + cc - dd;
  m[5] = 2 * (cd - ab);
  m[6] = 2 * (bd - ac);
  m[7] = 2 * (ab + cd);
  m[8] = aa - bb - cc + dd;
}

/* math.cpp:44-60: Rodrigues rotation about a unit axis */
void or_rotmat_axis(const double k[3], double t, double m[9]) {
  double jac[9];
  or_rotmat_axis_grad(k, t, m, jac);
}

/* math.cpp:62-94: rotation matrix and its derivative with respect to theta */
void or_rotmat_axis_grad(const double k[3], double t, double m[9], double j[9]) {
  double s = sin(t), c = cos(t), omc = 1 - c;
  double x = k[0], y = k[1], z = k[2];
  m[0] = c + x * x * omc;
  m[1] = x * y * omc - z * s;
  m[2] = y * s + x * z * omc;
  m[3] = z * s + x * y * omc;
  m[4] = c + y * y * omc;
  m[5] = y * z * omc - x * s;
  m[6] = x * z * omc - y * s;
  m[7] = x * s + y * z * omc;
  m[8] = c + z * z * omc;
  j[0] = -s + x * x * s;
  j[1] = -z * c + x * y * s;
  j[2] = y * c + x * z * s;
  j[3] = z * c + x * y * s;
  j[4] = -s + y * y * s;
  j[5] = -x * c + y * z * s;
  j[6] = -y * c + x * z * s;
  j[7] = x * c + y * z * s;
  j[8] = -s + z * z * s;
}

/* math.cpp:96-110: angle i-j-k by atan2(|v0 x v1|, v0.v1 + eps) and its
 * gradients on the outer atoms (the centre atom gets -(gi + gk)). */
double or_angle_grad(const double* pi, const double* pj, const double* pk, double* gi,
                     double* gk) {
  double v0[3], v1[3], cr[3];
  sub3(pj, pi, v0);
  sub3(pj, pk, v1);
  cross3(v0, v1, cr);
  double rp = norm3(cr);
  cross3(v0, cr, gi);
  cross3(cr, v1, gk);
  scale3(gi, -safe_div(1.0, sqsum3(v0) * rp));
  scale3(gk, -safe_div(1.0, sqsum3(v1) * rp));
  return atan2(rp, dot3(v0, v1) + EPS);
}

/* math.cpp:112-128 */
double or_dihedral(const double* pi, const double* pj, const double* pk, const double* pl) {
  double gi[3], gj[3], gk[3], gl[3];
  return or_dihedral_grad(pi, pj, pk, pl, gi, gj, gk, gl);
}

/* math.cpp:130-167: signed dihedral and its gradient on the four atoms */
double or_dihedral_grad(const double* pi, const double* pj, const double* pk,
                        const double* pl, double* gi, double* gj, double* gk, double* gl) {
  double rij[3], rkj[3], rkl[3], m[3], n[3], w[3];
  sub3(pj, pi, rij);
  sub3(pj, pk, rkj);
  sub3(pl, pk, rkl);
  cross3(rij, rkj, m);
  cross3(rkj, rkl, n);
  cross3(m, n, w);
  double wlen = norm3(w);
  double s = dot3(m, n);
  double phi = atan2(wlen, s + EPS);
  double ipr = dot3(rij, n);
  double theta = ipr < 0.0 ? phi : -phi;
  double nrkj2 = sqsum3(rkj);
  double nrkj_1 = safe_div(1.0, sqrt(nrkj2));
  double nrkj_2 = nrkj_1 * nrkj_1;
  double nrkj = nrkj2 * nrkj_1;
  double a = -nrkj / sqsum3(m);
  double b = nrkj / sqsum3(n);
  for (int d = 0; d < 3; ++d) {
    gi[d] = -a * m[d];
    gl[d] = -b * n[d];
  }
  double p = dot3(rij, rkj) * nrkj_2;
  double q = dot3(rkj, rkl) * nrkj_2;
  for (int d = 0; d < 3; ++d) {
    double sv = p * gi[d] - q * gl[d];
    gj[d] = -(gi[d] - sv);
    gk[d] = -(gl[d] + sv);
  }
  return theta;
}

/* math.h:198-208 (three-argument reversed ramp) */
double or_rev_linear(double v, double minv, double maxv) {
  if (v < minv) return 1.0;
  if (v <= maxv) return 1.0 - (v - minv) / (maxv - minv);
  return 0.0;
}

/* math.h:210-215 (ramp about a centre) */
double or_rev_linear_c(double v, double center, double minv, double maxv) {
  if (v < center) return or_rev_linear(center - v, minv, maxv);
  return or_rev_linear(v - center, minv, maxv);
}

/* ======================================================================== */
/* pair terms (pair.h:25-60).  rij points from i to j; gradients accumulate. */

double or_coul_pair(double qi, double qj, const double* rij, double d, double scale,
                    double* gi, double* gj) {
  double qq = qi * qj * scale;
  double de_dr_r = -safe_div(qq, d * d * d) * or_coulomb_factor;
  double g[3] = {rij[0] * de_dr_r, rij[1] * de_dr_r, rij[2] * de_dr_r};
  for (int k = 0; k < 3; ++k) {
    gi[k] -= g[k];
    gj[k] += g[k];
  }
  return safe_div(qq, d) * or_coulomb_factor;
}

double or_vdw_pair(double si, double ei, double sj, double ej, const double* rij, double d,
                   double scale, double* gi, double* gj) {
  double sig = (si + sj) * 0.5;
  double u6 = pow(safe_div(sig, d), 6.0);
  double u12 = u6 * u6;
  double eps4 = sqrt(ei * ej) * scale * 4.0;
  double irsqr = safe_div(1.0, d * d);
  double de_dr_r = (-12.0 * u12 + 6.0 * u6) * irsqr * eps4;
  double g[3] = {rij[0] * de_dr_r, rij[1] * de_dr_r, rij[2] * de_dr_r};
  for (int k = 0; k < 3; ++k) {
    gi[k] -= g[k];
    gj[k] += g[k];
  }
  return (u12 - u6) * eps4;
}

/* LJ parameters of an atom: ligands use the GAFF2 table by type
 * (interaction.cpp:119-121), receptors their own sigma/epsilon (:123-125). */
static inline void atom_lj(const pxd_mol* m, int i, double* s, double* e) {
  if (m->vdw_type) {
    int t = m->vdw_type[i];
    *s = or_gaff2_sigma[t];
    *e = or_gaff2_epsilon[t];
  } else {
    *s = m->sigma[i];
    *e = m->epsilon[i];
  }
}

/* ======================================================================== */
/* intra-molecular terms                                                     */

/* interaction.cpp:27-47 */
double or_bond_energy(const pxd_mol* m, const double* x, double* g) {
  double e = 0.0;
  for (int b = 0; b < m->n_bonds; ++b) {
    int i = m->bond_ij[2 * b], j = m->bond_ij[2 * b + 1];
    double d[3];
    sub3(x + 3 * j, x + 3 * i, d);
    double r = norm3(d);
    double dr = r - m->bond_r0[b];
    e += 0.5 * m->bond_k[b] * dr * dr;
    double de = m->bond_k[b] * (1.0 - safe_div(m->bond_r0[b], r));
    scale3(d, de);
    for (int k = 0; k < 3; ++k) {
      g[3 * i + k] -= d[k];
      g[3 * j + k] += d[k];
    }
  }
  return e;
}

/* interaction.cpp:49-74 */
double or_angle_energy(const pxd_mol* m, const double* x, double* g) {
  double e = 0.0;
  for (int a = 0; a < m->n_angles; ++a) {
    int i = m->angle_ijk[3 * a], j = m->angle_ijk[3 * a + 1], k = m->angle_ijk[3 * a + 2];
    double fi[3], fk[3];
    double th = or_angle_grad(x + 3 * i, x + 3 * j, x + 3 * k, fi, fk);
    double dl = th - m->angle_t0[a];
    e += 0.5 * m->angle_k[a] * dl * dl;
    double de = m->angle_k[a] * dl;
    for (int d = 0; d < 3; ++d) {
      fi[d] *= de;
      fk[d] *= de;
      g[3 * i + d] += fi[d];
      g[3 * k + d] += fk[d];
      g[3 * j + d] -= fi[d] + fk[d];
    }
  }
  return e;
}

/* interaction.cpp:76-117: impropers then propers, same functional form */
double or_dihedral_energy(const pxd_mol* m, const double* x, double* g) {
  double e = 0.0;
  for (int t = 0; t < m->n_dihedrals; ++t) {
    const int* id = m->dih_ijkl + 4 * t;
    double gi[3], gj[3], gk[3], gl[3];
    double th = or_dihedral_grad(x + 3 * id[0], x + 3 * id[1], x + 3 * id[2], x + 3 * id[3],
                                 gi, gj, gk, gl);
    for (int n = 0; n < PXD_DIH_TERMS; ++n) {
      int per = m->dih_n[PXD_DIH_TERMS * t + n];
      if (per == 0) break;
      double k = m->dih_k[PXD_DIH_TERMS * t + n];
      double dl = th * per - m->dih_phase[PXD_DIH_TERMS * t + n];
      e += k * (1.0 + cos(dl));
      double de = k * per * sin(dl);
      axpy3(g + 3 * id[0], -de, gi);
      axpy3(g + 3 * id[1], -de, gj);
      axpy3(g + 3 * id[2], -de, gk);
      axpy3(g + 3 * id[3], -de, gl);
    }
  }
  return e;
}

/* interaction.cpp:127-168: 1-4 pairs (Coulomb x0.833333333, LJ x0.5,
 * interaction.h:52-53) then all other pairs at scale 1 */
double or_nonbonded_energy(const pxd_mol* m, const double* x, double* g) {
  double e1 = 0.0, e2 = 0.0;
  for (int pass = 0; pass < 2; ++pass) {
    int n = pass == 0 ? m->n_pairs14 : m->n_pairs;
    const int* ij = pass == 0 ? m->pairs14 : m->pairs;
    double cs = pass == 0 ? 0.833333333 : 1.0;
    double vs = pass == 0 ? 0.5 : 1.0;
    for (int p = 0; p < n; ++p) {
      int i = ij[2 * p], j = ij[2 * p + 1];
      double r[3];
      sub3(x + 3 * j, x + 3 * i, r);
      double d = norm3(r);
      double si, ei, sj, ej;
      atom_lj(m, i, &si, &ei);
      atom_lj(m, j, &sj, &ej);
      e1 += or_coul_pair(m->charge[i], m->charge[j], r, d, cs, g + 3 * i, g + 3 * j);
      e2 += or_vdw_pair(si, ei, sj, ej, r, d, vs, g + 3 * i, g + 3 * j);
    }
  }
  return e1 + e2;
}

/* interaction.cpp:170-196 */
double or_intra_energy(const pxd_mol* m, const double* x, double* g) {
  double e = 0.0;
  e += or_bond_energy(m, x, g);
  e += or_angle_energy(m, x, g);
  e += or_dihedral_energy(m, x, g);
  e += or_nonbonded_energy(m, x, g);
  return e;
}

/* ======================================================================== */
/* grid maps                                                                 */

/* grid.cpp:164-194 for one probe point p: Coulomb for a unit probe charge and
 * LJ for each requested ligand type, each with the probe-position gradient. */
void or_grid_point(const pxd_mol* rec, const double* rx, const uint8_t* excluded,
                   const double* p, const int32_t* types, int32_t n_types, double* out) {
  int nm = n_types + 1;
  double tmp[3] = {0, 0, 0};
  memset(out, 0, sizeof(double) * 4 * nm);
  for (int a = 0; a < rec->n_atoms; ++a) {
    if (excluded && excluded[a]) continue;
    double v[3];
    sub3(rx + 3 * a, p, v);
    double d = norm3(v);
    out[0] += or_coul_pair(1.0, rec->charge[a], v, d, 1.0, out + 1, tmp);
    for (int t = 0; t < n_types; ++t) {
      int ty = types[t];
      double* o = out + 4 * (t + 1);
      o[0] += or_vdw_pair(or_gaff2_sigma[ty], or_gaff2_epsilon[ty], rec->sigma[a],
                          rec->epsilon[a], v, d, 1.0, o + 1, tmp);
    }
  }
}

typedef struct {
  const pxd_mol* rec;
  const double* rx;
  const uint8_t* excl;
  const pxd_box* box;
  const int32_t* types;
  int32_t n_types;
  double* maps;
  int i0, i1;
} grid_job;

static void* grid_worker(void* arg) {
  grid_job* jb = (grid_job*)arg;
  const pxd_box* b = jb->box;
  size_t P = (size_t)b->npoints[0] * b->npoints[1] * b->npoints[2];
  int nm = jb->n_types + 1;
  double* buf = (double*)malloc(sizeof(double) * 4 * nm);
  for (int i = jb->i0; i < jb->i1; ++i)
    for (int j = 0; j < b->npoints[1]; ++j)
      for (int k = 0; k < b->npoints[2]; ++k) {
        double p[3] = {b->init_xyz[0] + b->spacing * i, b->init_xyz[1] + b->spacing * j,
                       b->init_xyz[2] + b->spacing * k};
        or_grid_point(jb->rec, jb->rx, jb->excl, p, jb->types, jb->n_types, buf);
        size_t off = ((size_t)i * b->npoints[1] + j) * b->npoints[2] + k; /* grid.h:120-122 */
        for (int m = 0; m < nm; ++m) memcpy(jb->maps + 4 * (m * P + off), buf + 4 * m, 32);
      }
  free(buf);
  return NULL;
}

/* grid.cpp:115-148: x-slabs split over threads */
void or_grid_build(const pxd_mol* rec, const double* rx, const uint8_t* excl,
                   const pxd_box* box, const int32_t* types, int32_t n_types, double* maps,
                   int32_t nthreads) {
  int ni = box->npoints[0];
  if (nthreads < 1) nthreads = 1;
  if (nthreads > ni) nthreads = ni;
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
  grid_job* jobs = (grid_job*)malloc(sizeof(grid_job) * nthreads);
  int len = (ni + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t) {
    grid_job j = {rec, rx, excl, box, types, n_types, maps, t * len, (t + 1) * len};
    if (j.i1 > ni) j.i1 = ni;
    jobs[t] = j;
    pthread_create(&th[t], NULL, grid_worker, &jobs[t]);
  }
  for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  free(th);
  free(jobs);
}

/* grid.cpp:273-303: trilinear interpolation of value and stored gradient */
static double interp_map(const double* map, const pxd_box* b, const size_t* off,
                         const double* f, double* grad) {
  double x = f[0], y = f[1], z = f[2], mx = 1 - x, my = 1 - y, mz = 1 - z;
  double w[8] = {mx * my * mz, x * my * mz, mx * y * mz, x * y * mz,
                 mx * my * z,  x * my * z,  mx * y * z,  x * y * z};
  double acc[4] = {0, 0, 0, 0};
  for (int c = 0; c < 8; ++c)
    for (int q = 0; q < 4; ++q) acc[q] += map[4 * off[c] + q] * w[c];
  grad[0] = acc[1];
  grad[1] = acc[2];
  grad[2] = acc[3];
  (void)b;
  return acc[0];
}

/* grid.cpp:197-262 */
int32_t or_grid_get(const or_grid* g, int32_t type, double q, const double* xyz,
                    double* energy, double* grad) {
  const pxd_box* b = &g->box;
  int slot = -1;
  for (int t = 0; t < g->n_types; ++t)
    if (g->types[t] == type) slot = t + 1;
  if (slot < 0) return -1; /* caller error: type has no map */
  double f[3], gt[3], pen = 0.0;
  long lo[3];
  for (int d = 0; d < 3; ++d) {
    f[d] = (xyz[d] - b->init_xyz[d]) * b->spacing_inv;
    if (f[d] < 0.0) {
      lo[d] = 0;
      gt[d] = -1.0;
      pen -= f[d];
      f[d] = 0.0;
    } else if (f[d] < b->npoints_minus_one[d]) {
      lo[d] = (long)f[d];
      gt[d] = 0.0;
    } else {
      lo[d] = b->npoints_minus_one[d] - 1;
      gt[d] = 1.0;
      pen += f[d] - b->npoints_minus_one[d];
      f[d] = b->npoints_minus_one[d];
    }
    f[d] -= lo[d];
  }
  size_t ny = b->npoints[1], nz = b->npoints[2];
  size_t off[8];
  for (int c = 0; c < 8; ++c) {
    size_t i = lo[0] + (c & 1), j = lo[1] + ((c >> 1) & 1), k = lo[2] + ((c >> 2) & 1);
    off[c] = (i * ny + j) * nz + k;
  }
  size_t P = (size_t)b->npoints[0] * ny * nz;
  int hit = 1;
  if (pen > 0.0) {
    *energy += pen * b->spacing * g->slope;
    for (int d = 0; d < 3; ++d) grad[d] += gt[d] * g->slope;
    hit = 0;
  }
  double gg[3];
  *energy += interp_map(g->maps, b, off, f, gg) * q;
  if (hit) axpy3(grad, q, gg);
  *energy += interp_map(g->maps + 4 * P * slot, b, off, f, gg);
  if (hit) axpy3(grad, 1.0, gg);
  return hit;
}

/* ======================================================================== */
/* inter-molecular energy (interaction.cpp:202-257)                          */

double or_inter_energy(const pxd_mol* rec, const double* rx, const or_grid* grid,
                       const pxd_mol* lig, const double* lx, double* rg, double* lg) {
  double e = 0.0;
  for (int i = 0; i < lig->n_atoms; ++i) {
    double qi = lig->charge[i];
    int ty = lig->vdw_type[i];
    double si = or_gaff2_sigma[ty], ei = or_gaff2_epsilon[ty];
    if (grid) {
      or_grid_get(grid, ty, qi, lx + 3 * i, &e, lg + 3 * i);
      for (int x = 0; x < grid->n_excluded; ++x) {
        int j = grid->excluded[x];
        double r[3];
        sub3(rx + 3 * j, lx + 3 * i, r);
        double d = norm3(r);
        e += or_coul_pair(qi, rec->charge[j], r, d, 1.0, lg + 3 * i, rg + 3 * j);
        e += or_vdw_pair(si, ei, rec->sigma[j], rec->epsilon[j], r, d, 1.0, lg + 3 * i,
                         rg + 3 * j);
      }
    } else {
      for (int j = 0; j < rec->n_atoms; ++j) {
        double r[3];
        sub3(rx + 3 * j, lx + 3 * i, r);
        double d = norm3(r);
        e += or_coul_pair(qi, rec->charge[j], r, d, 1.0, lg + 3 * i, rg + 3 * j);
        e += or_vdw_pair(si, ei, rec->sigma[j], rec->epsilon[j], r, d, 1.0, lg + 3 * i,
                         rg + 3 * j);
      }
    }
  }
  return e;
}

/* ======================================================================== */
/* total energy with receptor torsions (interaction.cpp:271-350)             */

static int total_movable(const pxd_leaves* L) { return L ? L->atom_off[L->n_leaves] : 0; }

double or_system_energy(const or_system* s, const double* tors, const double* lx,
                        double* tg, double* lg) {
  const pxd_mol* rec = s->rec;
  int nr = rec->n_atoms, nl = s->lig->n_atoms;
  int nleaf = s->leaves ? s->leaves->n_leaves : 0;
  double* rx = (double*)malloc(sizeof(double) * 3 * nr);
  double* rg = (double*)calloc(3 * (size_t)nr, sizeof(double));
  double* jac = (double*)malloc(sizeof(double) * 3 * (total_movable(s->leaves) + 1));
  memcpy(rx, s->rec_xyz, sizeof(double) * 3 * nr);
  /* receptor_torsion2xyz (interaction.cpp:316-336): each leaf rotates its
   * movable atoms from the INITIAL coordinates and keeps dR/dtheta (x - o) */
  int off = 0;
  for (int k = 0; k < nleaf; ++k) {
    const pxd_leaves* L = s->leaves;
    double R[9], J[9];
    or_rotmat_axis_grad(L->axis + 3 * k, tors[k], R, J);
    for (int a = L->atom_off[k]; a < L->atom_off[k + 1]; ++a) {
      int j = L->atoms[a];
      double t[3];
      sub3(s->rec_xyz + 3 * j, L->origin + 3 * k, t);
      mv3(R, t, rx + 3 * j);
      for (int d = 0; d < 3; ++d) rx[3 * j + d] += L->origin[3 * k + d];
      mv3(J, t, jac + 3 * off);
      ++off;
    }
  }
  memset(lg, 0, sizeof(double) * 3 * nl);
  double e = 0.0;
  e += or_intra_energy(rec, rx, rg);
  e += or_intra_energy(s->lig, lx, lg);
  e += or_inter_energy(rec, rx, s->grid, s->lig, lx, rg, lg);
  /* receptor_xyz2torsion (interaction.cpp:338-350) */
  off = 0;
  for (int k = 0; k < nleaf; ++k) {
    const pxd_leaves* L = s->leaves;
    tg[k] = 0.0;
    for (int a = L->atom_off[k]; a < L->atom_off[k + 1]; ++a) {
      tg[k] += dot3(jac + 3 * off, rg + 3 * L->atoms[a]);
      ++off;
    }
  }
  free(rx);
  free(rg);
  free(jac);
  return e;
}

/* ======================================================================== */
/* L-BFGS with strong-Wolfe line search (optimizer.cpp:56-560)               */

static __thread double g_noise_rel = 0.0;
static __thread uint64_t g_noise_state = 0;

void or_set_eval_noise(double rel, uint64_t seed) {
  g_noise_rel = rel;
  g_noise_state = seed * 0x9E3779B97F4A7C15ull + 1;
}

/* splitmix64 -> uniform in [-1, 1) */
static double noise_u(void) {
  uint64_t z = (g_noise_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) / 4503599627370496.0 - 1.0;
}

typedef struct {
  const or_system* s;
  int n, nl3, nt;
  double *x0, *xt, *gtmp;
  int32_t* nevals;
  int mode;            /* 0 Cartesian, 1 rigid-body + torsion */
  const double* xyz0;  /* mode 1: input pose */
} closure;

/* the objective of either variable set at x */
static double clo_eval(closure* c, const double* x, double* g) {
  if (c->mode) return or_dof_energy(c->s, c->xyz0, x, g, NULL);
  return or_system_energy(c->s, x + c->nl3, x, g + c->nl3, g);
}

/* obj_func_closure::evaluate_direction (optimizer.cpp:155-167):
 * f and g at x0 + t d, x0 unchanged */
static double eval_dir(closure* c, double t, const double* d, double* g) {
  for (int i = 0; i < c->n; ++i) c->xt[i] = c->x0[i] + d[i] * t;
  double e = clo_eval(c, c->xt, g);
  if (g_noise_rel != 0.0) {
    e *= 1.0 + g_noise_rel * noise_u();
    for (int i = 0; i < c->n; ++i) g[i] *= 1.0 + g_noise_rel * noise_u();
  }
  ++*c->nevals;
  return e;
}

static double vdot(const double* a, const double* b, int n) {
  double v = 0.0;
  for (int i = 0; i < n; ++i) v += a[i] * b[i];
  return v;
}
static double vamax(const double* a, int n) {
  double v = 0.0;
  for (int i = 0; i < n; ++i) {
    double t = fabs(a[i]);
    if (t > v) v = t;
  }
  return v;
}

/* optimizer.cpp:205-233 */
static double cubic_interpolate(double x1, double f1, double g1, double x2, double f2,
                                double g2, double lo, double hi) {
  double d1 = g1 + g2 - 3.0 * (f1 - f2) / (x1 - x2);
  double d2sq = d1 * d1 - g1 * g2;
  if (d2sq >= 0.0) {
    double d2 = sqrt(d2sq), mp;
    if (x1 <= x2)
      mp = x2 - (x2 - x1) * ((g2 + d2 - d1) / (g2 - g1 + 2.0 * d2));
    else
      mp = x1 - (x1 - x2) * ((g1 + d2 - d1) / (g1 - g2 + 2.0 * d2));
    double t = mp > lo ? mp : lo;
    return t < hi ? t : hi;
  }
  return (lo + hi) / 2.0;
}

/* optimizer.cpp:235-427 */
static void strong_wolfe(closure* c, const pxd_lbfgs_params* p, double* t, const double* d,
                         double* f, double* g, double gtd, double* gwork) {
  int n = c->n;
  /* gwork holds 4 vectors: g_new, g_prev, bracket_g[0], bracket_g[1] */
  double* g_new = gwork;
  double* g_prev = gwork + n;
  double* bg[2] = {gwork + 2 * n, gwork + 3 * n};
  double br[2] = {0, 0}, bf[2] = {0, 0}, bgtd[2] = {0, 0};
  double f_new = eval_dir(c, *t, d, g_new);
  if (isnan(f_new)) {
    *f = f_new;
    return;
  }
  double gtd_new = vdot(g_new, d, n);
  if (fabs(gtd_new) > p->gtd_threshold) {
    *f = f_new;
    memcpy(g, g_new, sizeof(double) * n);
    return;
  }
  double t_prev = 0.0, f_prev = *f, gtd_prev = gtd;
  memcpy(g_prev, g, sizeof(double) * n);
  int done = 0, ls_iter = 0;
  while (ls_iter < p->max_ls) {
    if ((f_new > *f + p->c1 * *t * gtd) || (ls_iter > 1 && f_new >= f_prev)) {
      br[0] = t_prev; br[1] = *t; bf[0] = f_prev; bf[1] = f_new;
      memcpy(bg[0], g_prev, sizeof(double) * n);
      memcpy(bg[1], g_new, sizeof(double) * n);
      bgtd[0] = gtd_prev; bgtd[1] = gtd_new;
      break;
    }
    if (fabs(gtd_new) <= -p->c2 * gtd) {
      br[0] = *t; bf[0] = f_new;
      memcpy(bg[0], g_new, sizeof(double) * n);
      done = 1;
      break;
    }
    if (gtd_new >= 0) {
      br[0] = t_prev; br[1] = *t; bf[0] = f_prev; bf[1] = f_new;
      memcpy(bg[0], g_prev, sizeof(double) * n);
      memcpy(bg[1], g_new, sizeof(double) * n);
      bgtd[0] = gtd_prev; bgtd[1] = gtd_new;
      break;
    }
    double lo = *t + 0.01 * (*t - t_prev), hi = *t * 10.0, tmp = *t;
    *t = cubic_interpolate(t_prev, f_prev, gtd_prev, *t, f_new, gtd_new, lo, hi);
    t_prev = tmp;
    f_prev = f_new;
    memcpy(g_prev, g_new, sizeof(double) * n);
    gtd_prev = gtd_new;
    f_new = eval_dir(c, *t, d, g_new);
    gtd_new = vdot(g_new, d, n);
    ls_iter++;
  }
  double d_norm = 0.0;
  if (ls_iter == p->max_ls) {
    br[0] = 0; br[1] = *t; bf[0] = *f; bf[1] = f_new;
    memcpy(bg[0], g, sizeof(double) * n);
    memcpy(bg[1], g_new, sizeof(double) * n);
  } else {
    d_norm = vamax(d, n);
  }
  int insuf = 0, low = 0, high = 1;
  if (!done && bf[0] > bf[1]) {
    low = 1;
    high = 0;
  }
  while (!done && ls_iter < p->max_ls) {
    double lo, hi;
    if (br[0] > br[1]) { lo = br[1]; hi = br[0]; } else { lo = br[0]; hi = br[1]; }
    double eps = hi - lo;
    if (eps * d_norm < p->ls_tol_change) break;
    *t = cubic_interpolate(br[0], bf[0], bgtd[0], br[1], bf[1], bgtd[1], lo, hi);
    eps *= 0.1;
    double m1 = hi - *t, m2 = *t - lo;
    if ((m1 < m2 ? m1 : m2) < eps) {
      if (insuf || *t >= hi || *t <= lo) {
        if (fabs(*t - hi) < fabs(*t - lo))
          *t = hi - eps;
        else
          *t = lo + eps;
        insuf = 0;
      } else {
        insuf = 1;
      }
    } else {
      insuf = 0;
    }
    f_new = eval_dir(c, *t, d, g_new);
    gtd_new = vdot(g_new, d, n);
    ls_iter++;
    if (f_new > (*f + p->c1 * *t * gtd) || f_new >= bf[low]) {
      br[high] = *t; bf[high] = f_new;
      memcpy(bg[high], g_new, sizeof(double) * n);
      bgtd[high] = gtd_new;
      if (bf[0] > bf[1]) { low = 1; high = 0; } else { low = 0; high = 1; }
    } else {
      if (fabs(gtd_new) <= -p->c2 * gtd) {
        done = 1;
      } else if (gtd_new * (br[high] - br[low]) >= 0.0) {
        br[high] = br[low]; bf[high] = bf[low];
        memcpy(bg[high], bg[low], sizeof(double) * n);
        bgtd[high] = bgtd[low];
      }
      br[low] = *t; bf[low] = f_new;
      memcpy(bg[low], g_new, sizeof(double) * n);
      bgtd[low] = gtd_new;
    }
  }
  *t = br[low];
  *f = bf[low];
  memcpy(g, bg[low], sizeof(double) * n);
}

/* lbfgs_step::apply (optimizer.cpp:429-560).  The flat variable is
 * [ligand xyz (3N) | receptor torsions (T)] (optimizer.cpp:56-76). */
int32_t or_lbfgs(const or_system* s, const pxd_lbfgs_params* p, double* lx, double* tors,
                 double* energy, int32_t* nevals) {
  const int mode = p->dof_mode ? 1 : 0;
  if (mode && !s->tree) return 0;
  int nl3 = mode ? 6 + s->tree->n_torsions : 3 * s->lig->n_atoms;
  int nt = s->leaves ? s->leaves->n_leaves : 0;
  int n = nl3 + nt, m = p->history;
  double* buf = (double*)calloc((size_t)n * (14 + 2 * (size_t)m) + 4 * (size_t)m, sizeof(double));
  double *x = buf, *g = x + n, *gp = g + n, *y = gp + n, *d = y + n, *sv = d + n, *q = sv + n;
  double *xt = q + n, *gw = xt + n; /* gw: 4n */
  double* dirs = gw + 4 * n;        /* m x n (y history) */
  double* stps = dirs + (size_t)m * n; /* m x n (s history) */
  double* ro = stps + (size_t)m * n;
  double* al = ro + m;
  int head = 0, num_old = 0; /* ring buffer of history */
  double* xyz0 = NULL;
  if (mode) { /* variables start at zero: the input pose itself */
    xyz0 = (double*)malloc(sizeof(double) * 3 * s->lig->n_atoms);
    memcpy(xyz0, lx, sizeof(double) * 3 * s->lig->n_atoms);
  } else {
    memcpy(x, lx, sizeof(double) * nl3);
  }
  if (nt) memcpy(x + nl3, tors, sizeof(double) * nt);
  *nevals = 0;
  closure c = {s, n, nl3, nt, x, xt, NULL, nevals, mode, xyz0};
  double loss = clo_eval(&c, x, g);
  *energy = loss;
  *nevals = 1;
  if (vamax(g, n) < p->tol_grad) {
    free(buf);
    free(xyz0);
    return 1;
  }
  int high_change = 0, converged = 0, updated = 0, n_iter = 0;
  double H = 1.0, t = 0.0, prev_loss;
  while (n_iter < p->max_iter && *nevals < p->max_eval) {
    n_iter++;
    if (n_iter > 1) {
      for (int i = 0; i < n; ++i) y[i] = g[i] - gp[i];
      high_change = vamax(y, n) > 1e8;
    }
    if (n_iter == 1 || high_change) {
      for (int i = 0; i < n; ++i) d[i] = -g[i];
      num_old = 0;
      head = 0;
      H = 1.0;
    } else {
      for (int i = 0; i < n; ++i) sv[i] = d[i] * t;
      double ys = vdot(y, sv, n);
      if (ys > 1e-10) {
        int slot;
        if (num_old == m) { /* drop the oldest (optimizer.cpp:474-478) */
          slot = head;
          head = (head + 1) % m;
        } else {
          slot = (head + num_old) % m;
          num_old++;
        }
        H = ys / vdot(y, y, n);
        memcpy(dirs + (size_t)slot * n, y, sizeof(double) * n);
        memcpy(stps + (size_t)slot * n, sv, sizeof(double) * n);
        ro[slot] = 1.0 / ys;
      }
      for (int i = 0; i < n; ++i) q[i] = -g[i];
      for (int h = num_old - 1; h >= 0; --h) {
        int slot = (head + h) % m;
        al[h] = vdot(stps + (size_t)slot * n, q, n) * ro[slot];
        for (int i = 0; i < n; ++i) q[i] -= al[h] * dirs[(size_t)slot * n + i];
      }
      for (int i = 0; i < n; ++i) d[i] = q[i] * H;
      for (int h = 0; h < num_old; ++h) {
        int slot = (head + h) % m;
        double be = vdot(dirs + (size_t)slot * n, d, n) * ro[slot];
        for (int i = 0; i < n; ++i) d[i] += stps[(size_t)slot * n + i] * (al[h] - be);
      }
    }
    memcpy(gp, g, sizeof(double) * n);
    prev_loss = loss;
    if (n_iter == 1 || high_change) {
      double l1 = 0.0;
      for (int i = 0; i < n; ++i) l1 += fabs(g[i]);
      double inv = 1.0 / l1;
      t = (1.0 < inv ? 1.0 : inv) * p->lr * 0.1;
    } else {
      t = p->lr;
    }
    double gtd = vdot(g, d, n);
    if (gtd > -p->tol_change) {
      converged = 1;
      break;
    }
    strong_wolfe(&c, p, &t, d, &loss, g, gtd, gw);
    if (isnan(loss)) break;
    for (int i = 0; i < n; ++i) x[i] += d[i] * t;
    updated = 1;
    if (vamax(g, n) <= p->tol_grad) {
      converged = 1;
      break;
    }
    double smax = 0.0;
    for (int i = 0; i < n; ++i) {
      double v = fabs(d[i] * t);
      if (v > smax) smax = v;
    }
    if (smax <= p->tol_change) {
      converged = 1;
      break;
    }
    if (fabs(loss - prev_loss) < p->tol_change) {
      converged = 1;
      break;
    }
  }
  if (updated) {
    *energy = clo_eval(&c, x, g);
    ++*nevals;
  }
  if (mode)
    or_dof_energy(s, xyz0, x, g, lx); /* the final pose (energy already reported) */
  else
    memcpy(lx, x, sizeof(double) * nl3);
  if (nt) memcpy(tors, x + nl3, sizeof(double) * nt);
  free(buf);
  free(xyz0);
  return converged;
}

/* ---- rigid-body + torsion variables ---- */

void or_dof_grad(const pxd_frag_tree* t, int32_t n, const double* x, const double* g,
                 const double center[3], double* out) {
  double tq[PXD_MAX_FRAGMENTS][3];
  memset(tq, 0, sizeof(tq));
  out[0] = out[1] = out[2] = out[3] = out[4] = out[5] = 0.0;
  for (int p = 0; p < n; ++p) {
    const int a = t->perm[p];
    const double* xa = x + 3 * a;
    const double* ga = g + 3 * a;
    double r[3], c[3];
    sub3(xa, center, r);
    cross3(r, ga, c);
    for (int d = 0; d < 3; ++d) {
      out[d] += ga[d];
      out[3 + d] += c[d];
    }
    int f = 0;
    while (f < t->n_frags && !(p >= t->frag_start[f] && p < t->frag_end[f])) ++f;
    /* every torsion on the path to the root moves this atom */
    for (; f < t->n_frags && t->frag_parent[f] >= 0; f = t->frag_parent[f]) {
      const double* om = x + 3 * t->perm[t->frag_bond[2 * f + 1]];
      sub3(xa, om, r);
      cross3(r, ga, c);
      axpy3(tq[f], 1.0, c);
    }
  }
  for (int k = 0; k < t->n_torsions; ++k) {
    int f = t->tors_frag[k];
    double ax[3];
    sub3(x + 3 * t->perm[t->frag_bond[2 * f + 1]], x + 3 * t->perm[t->frag_bond[2 * f]], ax);
    scale3(ax, 1.0 / norm3(ax));
    out[6 + k] = dot3(ax, tq[f]);
  }
}

/* left Jacobian of SO(3) at w: exp([w + e]x) = exp([J e]x) exp([w]x) + O(e^2) */
static void so3_left_jacobian(const double* w, double* J) {
  double th2 = sqsum3(w), th = sqrt(th2), a, b;
  if (th < 1e-4) {
    a = 0.5 - th2 / 24.0;
    b = 1.0 / 6.0 - th2 / 120.0;
  } else {
    a = (1.0 - cos(th)) / th2;
    b = (th - sin(th)) / (th2 * th);
  }
  double K[9] = {0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0}, K2[9];
  or_matmul3(K, K, K2);
  for (int i = 0; i < 9; ++i) J[i] = (i % 4 == 0 ? 1.0 : 0.0) + a * K[i] + b * K2[i];
}

double or_dof_energy(const or_system* s, const double* xyz0, const double* x, double* grad,
                     double* pose_out) {
  const pxd_frag_tree* t = s->tree;
  const int n = s->lig->n_atoms, nlt = t->n_torsions;
  double c0[3] = {0, 0, 0}, cen[3], R[9];
  for (int a = 0; a < n; ++a) axpy3(c0, 1.0, xyz0 + 3 * a);
  scale3(c0, 1.0 / n);
  for (int d = 0; d < 3; ++d) cen[d] = c0[d] + x[d];
  double th = norm3(x + 3);
  if (th > 0.0) {
    double k[3] = {x[3] / th, x[4] / th, x[5] / th};
    or_rotmat_axis(k, th, R);
  } else {
    double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    memcpy(R, I, sizeof(I));
  }
  double* pose = (double*)malloc(sizeof(double) * 6 * n);
  double* gc = pose + 3 * n;
  or_ligand_apply(t, n, xyz0, cen, R, x + 6, pose);
  int nt = s->leaves ? s->leaves->n_leaves : 0;
  double e = or_system_energy(s, x + 6 + nlt, pose, grad + 6 + nlt, gc);
  double dof[6 + PXD_MAX_FRAGMENTS], J[9], tau[3];
  or_dof_grad(t, n, pose, gc, cen, dof);
  so3_left_jacobian(x + 3, J);
  for (int i = 0; i < 3; ++i) tau[i] = dof[3 + i];
  for (int i = 0; i < 3; ++i) {
    grad[i] = dof[i];
    grad[3 + i] = J[i] * tau[0] + J[3 + i] * tau[1] + J[6 + i] * tau[2];  /* J^T tau */
  }
  for (int k = 0; k < nlt; ++k) grad[6 + k] = dof[6 + k];
  (void)nt;
  if (pose_out) memcpy(pose_out, pose, sizeof(double) * 3 * n);
  free(pose);
  return e;
}

/* ---- threaded batch drivers (one pose per task, entrance.cpp:371-380) ---- */

typedef struct {
  const or_system* systems;
  const pxd_lbfgs_params* p;
  int32_t n_poses;
  const int32_t *pose_lig, *pose_atom_off;
  double *xyz, *tors, *energy, *lig_grad, *tors_grad;
  int32_t n_leaves;
  int32_t *nevals, *converged;
  int mode; /* 0 score, 1 lbfgs */
  volatile int next;
  pthread_mutex_t mu;
} many_job;

static void* many_worker(void* arg) {
  many_job* j = (many_job*)arg;
  for (;;) {
    pthread_mutex_lock(&j->mu);
    int pi = j->next++;
    pthread_mutex_unlock(&j->mu);
    if (pi >= j->n_poses) break;
    const or_system* s = j->systems + j->pose_lig[pi];
    double* x = j->xyz + 3 * (size_t)j->pose_atom_off[pi];
    double* t = j->tors ? j->tors + (size_t)pi * j->n_leaves : NULL;
    double zero[PXD_MAX_LEAVES] = {0};
    if (j->mode == 0) {
      double tg[PXD_MAX_LEAVES];
      double* g = j->lig_grad + 3 * (size_t)j->pose_atom_off[pi];
      j->energy[pi] = or_system_energy(s, t ? t : zero, x, tg, g);
      if (j->tors_grad)
        memcpy(j->tors_grad + (size_t)pi * j->n_leaves, tg, sizeof(double) * j->n_leaves);
    } else {
      j->converged[pi] = or_lbfgs(s, j->p, x, t ? t : zero, j->energy + pi, j->nevals + pi);
    }
  }
  return NULL;
}

static void run_many(many_job* j, int nthreads) {
  if (nthreads < 1) nthreads = 1;
  pthread_mutex_init(&j->mu, NULL);
  j->next = 0;
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nthreads);
  for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, many_worker, j);
  for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
  free(th);
  pthread_mutex_destroy(&j->mu);
}

void or_score_many(const or_system* systems, int32_t n_poses, const int32_t* pose_lig,
                   const int32_t* pose_atom_off, const double* xyz, const double* tors,
                   int32_t n_leaves, double* energy, double* lig_grad, double* tors_grad,
                   int32_t nthreads) {
  many_job j;
  memset(&j, 0, sizeof(j));
  j.systems = systems;
  j.n_poses = n_poses;
  j.pose_lig = pose_lig;
  j.pose_atom_off = pose_atom_off;
  j.xyz = (double*)xyz;
  j.tors = (double*)tors;
  j.n_leaves = n_leaves;
  j.energy = energy;
  j.lig_grad = lig_grad;
  j.tors_grad = tors_grad;
  j.mode = 0;
  run_many(&j, nthreads);
}

void or_lbfgs_many(const or_system* systems, const pxd_lbfgs_params* p, int32_t n_poses,
                   const int32_t* pose_lig, const int32_t* pose_atom_off, double* xyz,
                   double* tors, int32_t n_leaves, double* energy, int32_t* nevals,
                   int32_t* converged, int32_t nthreads) {
  many_job j;
  memset(&j, 0, sizeof(j));
  j.systems = systems;
  j.p = p;
  j.n_poses = n_poses;
  j.pose_lig = pose_lig;
  j.pose_atom_off = pose_atom_off;
  j.xyz = xyz;
  j.tors = tors;
  j.n_leaves = n_leaves;
  j.energy = energy;
  j.nevals = nevals;
  j.converged = converged;
  j.mode = 1;
  run_many(&j, nthreads);
}

/* ======================================================================== */
/* conformations                                                             */

/* torsional_receptor::apply_parameters (data.cpp:366-385) */
void or_receptor_apply(const pxd_leaves* L, int32_t n, const double* xyz, const double* tors,
                       double* out) {
  memcpy(out, xyz, sizeof(double) * 3 * n);
  for (int k = 0; L && k < L->n_leaves; ++k) {
    double R[9];
    or_rotmat_axis(L->axis + 3 * k, tors[k], R);
    for (int a = L->atom_off[k]; a < L->atom_off[k + 1]; ++a) {
      double* p = out + 3 * L->atoms[a];
      double t[3];
      sub3(p, L->origin + 3 * k, t);
      mv3(R, t, p);
      for (int d = 0; d < 3; ++d) p[d] += L->origin[3 * k + d];
    }
  }
}

/* free_ligand::apply_parameters (data.cpp:522-641), restated per fragment:
 * with local coordinates l (root atoms and first-level bond atoms relative
 * to the old geometric centre, other atoms relative to their fragment's bond
 * atom), M_root = rot and for a fragment f with parent p:
 *   axis'_f = M_p axis_f,  M_f = R(axis'_f, theta_f) M_p,
 *   y[mine_f] = M_p l[mine_f] + O_p,  y_i = M_f l_i + O_f (i in f, i != mine_f)
 * where O_root = new centre and O_f = y[mine_f]. */
void or_ligand_apply(const pxd_frag_tree* t, int32_t n, const double* xyz,
                     const double center[3], const double rot[9], const double* tors,
                     double* out) {
  double* cp = (double*)malloc(sizeof(double) * 3 * n);
  double* loc = (double*)malloc(sizeof(double) * 3 * n);
  double axes[PXD_MAX_FRAGMENTS][3], M[PXD_MAX_FRAGMENTS][9], O[PXD_MAX_FRAGMENTS][3];
  double theta[PXD_MAX_FRAGMENTS];
  for (int p = 0; p < n; ++p) memcpy(cp + 3 * p, xyz + 3 * t->perm[p], 24); /* data.cpp:522-527 */
  /* rotatable-bond axes in the input frame (data.cpp:530-536) */
  for (int f = 0; f < t->n_frags; ++f) {
    theta[f] = 0.0;
    if (t->frag_parent[f] < 0) continue;
    sub3(cp + 3 * t->frag_bond[2 * f + 1], cp + 3 * t->frag_bond[2 * f], axes[f]);
    double nn = norm3(axes[f]);
    scale3(axes[f], 1.0 / nn);
  }
  for (int k = 0; k < t->n_torsions; ++k) theta[t->tors_frag[k]] = tors[k];
  /* geometric centre of the input (data.cpp:606) */
  double c[3] = {0, 0, 0};
  for (int p = 0; p < n; ++p) axpy3(c, 1.0, cp + 3 * p);
  scale3(c, 1.0 / n);
  /* local coordinates (data.cpp:538-554, 605-615) */
  for (int f = 0; f < t->n_frags; ++f) {
    int par = t->frag_parent[f];
    if (par < 0) {
      for (int p = t->frag_start[f]; p < t->frag_end[f]; ++p) sub3(cp + 3 * p, c, loc + 3 * p);
      continue;
    }
    int mine = t->frag_bond[2 * f + 1];
    for (int p = t->frag_start[f]; p < t->frag_end[f]; ++p)
      if (p != mine) sub3(cp + 3 * p, cp + 3 * mine, loc + 3 * p);
    const double* ref = t->frag_parent[par] < 0 ? c : cp + 3 * t->frag_bond[2 * par + 1];
    sub3(cp + 3 * mine, ref, loc + 3 * mine);
  }
  /* roots (data.cpp:617-631) */
  for (int f = 0; f < t->n_frags; ++f) {
    if (t->frag_parent[f] >= 0) continue;
    memcpy(M[f], rot, sizeof(double) * 9);
    memcpy(O[f], center, 24);
    for (int p = t->frag_start[f]; p < t->frag_end[f]; ++p) {
      mv3(rot, loc + 3 * p, out + 3 * p);
      axpy3(out + 3 * p, 1.0, center);
    }
  }
  /* children in pre-order (data.cpp:556-590) */
  for (int k = 0; k < t->n_torsions; ++k) {
    int f = t->tors_frag[k], par = t->frag_parent[f];
    int mine = t->frag_bond[2 * f + 1];
    double ax[3], R[9];
    mv3(M[par], axes[f], ax);
    mv3(M[par], loc + 3 * mine, out + 3 * mine);
    axpy3(out + 3 * mine, 1.0, O[par]);
    or_rotmat_axis(ax, theta[f], R);
    or_matmul3(R, M[par], M[f]);
    memcpy(O[f], out + 3 * mine, 24);
    for (int p = t->frag_start[f]; p < t->frag_end[f]; ++p) {
      if (p == mine) continue;
      mv3(M[f], loc + 3 * p, out + 3 * p);
      axpy3(out + 3 * p, 1.0, O[f]);
    }
  }
  /* un-permute (data.cpp:640) */
  memcpy(cp, out, sizeof(double) * 3 * n);
  for (int p = 0; p < n; ++p) memcpy(out + 3 * t->perm[p], cp + 3 * p, 24);
  free(cp);
  free(loc);
}

/* ======================================================================== */
/* pscore leaves (scorer.cpp:128-613)                                         */

static double dist3(const double* a, const double* b) {
  double v[3];
  sub3(a, b, v);
  return norm3(v);
}

/* scorer.cpp:128-137 */
static void ring_centroid(const double* x, const int* ids, int size, double* out) {
  out[0] = out[1] = out[2] = 0.0;
  for (int i = 0; i < size; ++i) axpy3(out, 1.0, x + 3 * ids[i]);
  out[0] /= size;
  out[1] /= size;
  out[2] /= size;
}
/* scorer.cpp:151-162: normalised (p2 - p1) x (p2 - p3) */
static void ring_normal(const double* x, const int* ids, double* out) {
  double e1[3], e2[3];
  sub3(x + 3 * ids[1], x + 3 * ids[0], e1);
  sub3(x + 3 * ids[1], x + 3 * ids[2], e2);
  cross3(e1, e2, out);
  double nn = norm3(out);
  out[0] /= nn;
  out[1] /= nn;
  out[2] /= nn;
}

/* scorer.cpp:222-252 */
double or_pscore_vdw(const pxd_score_params* sp, const pxd_mol* rec, const double* rx,
                     const pxd_mol* lig, const double* lx) {
  double e = 0.0;
  for (int i = 0; i < lig->n_atoms; ++i) {
    int ty = lig->vdw_type[i];
    double s1 = or_gaff2_sigma[ty] * sp->vdw_sigma_scale[ty];
    double e1 = or_gaff2_epsilon[ty] * sp->vdw_epsilon_scale[ty];
    for (int j = 0; j < rec->n_atoms; ++j) {
      double d = dist3(lx + 3 * i, rx + 3 * j);
      double sig = (s1 + rec->sigma[j]) * 0.5;
      double u6 = pow(safe_div(sig, d), 6.0);
      double ee = sqrt(e1 * rec->epsilon[j]) * 4.0;
      e += clampd((u6 * u6 - u6) * ee, -sp->vdw_cutoff, sp->vdw_cutoff);
    }
  }
  return e * sp->vdw_scale;
}

/* scorer.cpp:254-276 */
static double ps_coulomb(const pxd_mol* rec, const double* rx, const pxd_mol* lig,
                         const double* lx) {
  double e = 0.0;
  for (int i = 0; i < lig->n_atoms; ++i)
    for (int j = 0; j < rec->n_atoms; ++j)
      e += safe_div(lig->charge[i] * rec->charge[j], dist3(lx + 3 * i, rx + 3 * j)) *
           or_coulomb_factor;
  return e;
}

/* scorer.cpp:278-311 */
static double ps_hydrophobic(const pxd_score_params* sp, const pxd_mol* rec, const double* rx,
                             const pxd_mol* lig, const double* lx) {
  double e = 0.0;
  for (int a = 0; a < lig->n_hydrophobic; ++a) {
    int i = lig->hydrophobic[a];
    double s1 = or_gaff2_sigma[lig->vdw_type[i]];
    for (int b = 0; b < rec->n_hydrophobic; ++b) {
      int j = rec->hydrophobic[b];
      double r1 = (s1 + rec->sigma[j]) * sp->hydro_scale_factor + sp->hydro_dist_min;
      double r2 = r1 + sp->hydro_dist_max;
      e -= or_rev_linear(dist3(lx + 3 * i, rx + 3 * j), r1, r2);
    }
  }
  return e * sp->hydro_scale;
}

/* scorer.cpp:341-368 */
static double ps_ionic_batch(const double* cx, const int* cat, int nc, const double* ax,
                             const int* an, int na, const double* range) {
  double e = 0.0;
  for (int i = 0; i < nc; ++i)
    for (int j = 0; j < na; ++j)
      e -= or_rev_linear(dist3(cx + 3 * cat[i], ax + 3 * an[j]), range[0], range[1]);
  return e;
}

/* scorer.cpp:322-339 */
static double ps_ionic(const pxd_score_params* sp, const pxd_mol* rec, const double* rx,
                       const pxd_mol* lig, const double* lx) {
  double e = 0.0;
  e += sp->ionic_lc_coef * ps_ionic_batch(lx, lig->cation, lig->n_cation, rx, rec->anion,
                                          rec->n_anion, sp->ionic_lc);
  e += sp->ionic_pc_coef * ps_ionic_batch(rx, rec->cation, rec->n_cation, lx, lig->anion,
                                          lig->n_anion, sp->ionic_pc);
  return e;
}

/* scorer.cpp:447-462 */
static double ps_cation_pi_batch(const double* cx, const int* cat, int nc,
                                 const double* cents, int nr, const double* range) {
  double e = 0.0;
  for (int r = 0; r < nr; ++r)
    for (int j = 0; j < nc; ++j)
      e -= or_rev_linear(dist3(cents + 3 * r, cx + 3 * cat[j]), range[0], range[1]);
  return e;
}

typedef struct {
  double *rc5, *rc6, *lc5, *lc6, *rn5, *rn6, *ln5, *ln6;
} ring_memo;

/* precalculate ring centroids and normals (scorer.cpp:190-203) */
static void ring_memo_fill(ring_memo* m, const pxd_mol* rec, const double* rx,
                           const pxd_mol* lig, const double* lx) {
  int n = 3 * (rec->n_ring5 + rec->n_ring6 + lig->n_ring5 + lig->n_ring6) + 3;
  double* b = (double*)malloc(sizeof(double) * 2 * n);
  m->rc5 = b;
  m->rc6 = m->rc5 + 3 * rec->n_ring5;
  m->lc5 = m->rc6 + 3 * rec->n_ring6;
  m->lc6 = m->lc5 + 3 * lig->n_ring5;
  m->rn5 = b + n;
  m->rn6 = m->rn5 + 3 * rec->n_ring5;
  m->ln5 = m->rn6 + 3 * rec->n_ring6;
  m->ln6 = m->ln5 + 3 * lig->n_ring5;
  for (int r = 0; r < rec->n_ring5; ++r) {
    ring_centroid(rx, rec->ring5 + 5 * r, 5, m->rc5 + 3 * r);
    ring_normal(rx, rec->ring5 + 5 * r, m->rn5 + 3 * r);
  }
  for (int r = 0; r < rec->n_ring6; ++r) {
    ring_centroid(rx, rec->ring6 + 6 * r, 6, m->rc6 + 3 * r);
    ring_normal(rx, rec->ring6 + 6 * r, m->rn6 + 3 * r);
  }
  for (int r = 0; r < lig->n_ring5; ++r) {
    ring_centroid(lx, lig->ring5 + 5 * r, 5, m->lc5 + 3 * r);
    ring_normal(lx, lig->ring5 + 5 * r, m->ln5 + 3 * r);
  }
  for (int r = 0; r < lig->n_ring6; ++r) {
    ring_centroid(lx, lig->ring6 + 6 * r, 6, m->lc6 + 3 * r);
    ring_normal(lx, lig->ring6 + 6 * r, m->ln6 + 3 * r);
  }
}

/* scorer.cpp:383-427 (memo branch) */
static double ps_cation_pi(const pxd_score_params* sp, const pxd_mol* rec, const double* rx,
                           const pxd_mol* lig, const double* lx, const ring_memo* m) {
  double e = 0.0;
  e += sp->pication_sub_coef[0] * ps_cation_pi_batch(lx, lig->cation, lig->n_cation, m->rc5,
                                                     rec->n_ring5, sp->pication_range[0]);
  e += sp->pication_sub_coef[1] * ps_cation_pi_batch(lx, lig->cation, lig->n_cation, m->rc6,
                                                     rec->n_ring6, sp->pication_range[1]);
  e += sp->pication_sub_coef[2] * ps_cation_pi_batch(rx, rec->cation, rec->n_cation, m->lc5,
                                                     lig->n_ring5, sp->pication_range[2]);
  e += sp->pication_sub_coef[3] * ps_cation_pi_batch(rx, rec->cation, rec->n_cation, m->lc6,
                                                     lig->n_ring6, sp->pication_range[3]);
  return e;
}

/* scorer.cpp:535-577 */
static double ps_hbond_batch(const double* dx, const int* don, int nd, const double* ax,
                             const int* acc, int na, const double* s) {
  /* s = (dist_min, dist_max, angle_min, angle_max, dist_center, angle_center) */
  double e = 0.0;
  for (int i = 0; i < nd; ++i) {
    const double* dp = dx + 3 * don[2 * i];
    const double* hp = dx + 3 * don[2 * i + 1];
    for (int j = 0; j < na; ++j) {
      const double* ap = ax + 3 * acc[j];
      double dr = or_rev_linear_c(dist3(hp, ap), s[4], s[0], s[1]);
      if (dr < EPS) continue;
      double v1[3], v2[3];
      sub3(dp, hp, v1);
      sub3(ap, hp, v2);
      double dpv = dot3(v1, v2), ll = norm3(v1) * norm3(v2);
      double ang = acos(clampd(safe_div(dpv, ll), -1.0, 1.0)) / PI * 180.0;
      e -= dr * or_rev_linear_c(ang, s[5], s[2], s[3]);
    }
  }
  return e;
}

/* scorer.cpp:486-533 */
static double ps_hbond(const pxd_score_params* sp, const pxd_mol* rec, const double* rx,
                       const pxd_mol* lig, const double* lx) {
  const double* w = sp->hbond_sub_coef;
  double e = 0.0;
  e += w[0] * ps_hbond_batch(lx, lig->don_charged, lig->n_don_charged, rx, rec->acc_charged,
                             rec->n_acc_charged, sp->hbond_inter[0]);
  e += w[1] * ps_hbond_batch(lx, lig->don_charged, lig->n_don_charged, rx, rec->acc_neutral,
                             rec->n_acc_neutral, sp->hbond_inter[1]);
  e += w[2] * ps_hbond_batch(lx, lig->don_neutral, lig->n_don_neutral, rx, rec->acc_charged,
                             rec->n_acc_charged, sp->hbond_inter[2]);
  e += w[3] * ps_hbond_batch(lx, lig->don_neutral, lig->n_don_neutral, rx, rec->acc_neutral,
                             rec->n_acc_neutral, sp->hbond_inter[3]);
  e += w[4] * ps_hbond_batch(rx, rec->don_charged, rec->n_don_charged, lx, lig->acc_charged,
                             lig->n_acc_charged, sp->hbond_inter[4]);
  e += w[5] * ps_hbond_batch(rx, rec->don_charged, rec->n_don_charged, lx, lig->acc_neutral,
                             lig->n_acc_neutral, sp->hbond_inter[5]);
  e += w[6] * ps_hbond_batch(rx, rec->don_neutral, rec->n_don_neutral, lx, lig->acc_charged,
                             lig->n_acc_charged, sp->hbond_inter[6]);
  e += w[7] * ps_hbond_batch(rx, rec->don_neutral, rec->n_don_neutral, lx, lig->acc_neutral,
                             lig->n_acc_neutral, sp->hbond_inter[7]);
  return e * sp->hbond_scale;
}

/* scorer.cpp:579-613: periods 1..4, phases {0, pi, 0, pi} */
static double ps_tstrain(const pxd_mol* lig, const double* lx) {
  static const double ph[4] = {0.0, PI, 0.0, PI};
  double score = 0.0;
  for (int p = 0; p < lig->n_tstrain; ++p) {
    const double* cf = lig->tstrain_coef + 4 * p;
    double tmp = 0.0;
    int j = 0, cnt = lig->tstrain_match_off[p + 1] - lig->tstrain_match_off[p];
    for (j = 0; j < cnt; ++j) {
      const int* a = lig->tstrain_ijkl + 4 * (lig->tstrain_match_off[p] + j);
      if (a[0] == a[1]) break;
      double th = or_dihedral(lx + 3 * a[0], lx + 3 * a[1], lx + 3 * a[2], lx + 3 * a[3]);
      for (int n = 0; n < 4; ++n) tmp += (1.0 + cos(th * (n + 1) - ph[n])) * cf[n];
    }
    if (j > 0) score += tmp / j;
  }
  return score;
}

/* scorer.cpp:615-636 */
static double proj_dist(const double* n1, const double* c1, const double* c2) {
  double t[3];
  sub3(c2, c1, t);
  double dp = dot3(n1, t);
  return sqrt(sqsum3(t) - dp * dp);
}

/* scorer.cpp:697-724: s = (dist_min, dist_max, angle_min, angle_max, sigma_min, sigma_max) */
static double ps_stack_batch(double center, const double* s, const double* rc,
                             const double* rn, int nr, const double* lc, const double* ln,
                             int nl) {
  double e = 0.0;
  for (int i = 0; i < nr; ++i)
    for (int j = 0; j < nl; ++j) {
      double d = dist3(rc + 3 * i, lc + 3 * j);
      if (d > s[1]) continue;
      d = or_rev_linear(d, s[0], s[1]);
      double s1 = proj_dist(rn + 3 * i, rc + 3 * i, lc + 3 * j);
      double s2 = proj_dist(ln + 3 * j, lc + 3 * j, rc + 3 * i);
      double sg = s1 < s2 ? s1 : s2;
      if (sg > s[5]) continue;
      sg = or_rev_linear(sg, s[4], s[5]);
      double th = acos(clampd(dot3(rn + 3 * i, ln + 3 * j), -1.0, 1.0)) / PI * 180.0;
      th = th < 180.0 - th ? th : 180.0 - th;
      th = or_rev_linear_c(th, center, s[2], s[3]);
      e -= d * sg * th;
    }
  return e;
}

/* scorer.cpp:651-695 (memo branch) */
static double ps_stack(double center, const double* w, const double (*inter)[6],
                       const pxd_mol* rec, const pxd_mol* lig, const ring_memo* m) {
  double e = 0.0;
  e += w[0] * ps_stack_batch(center, inter[0], m->rc5, m->rn5, rec->n_ring5, m->lc5, m->ln5,
                             lig->n_ring5);
  e += w[1] * ps_stack_batch(center, inter[1], m->rc6, m->rn6, rec->n_ring6, m->lc5, m->ln5,
                             lig->n_ring5);
  e += w[2] * ps_stack_batch(center, inter[2], m->rc5, m->rn5, rec->n_ring5, m->lc6, m->ln6,
                             lig->n_ring6);
  e += w[3] * ps_stack_batch(center, inter[3], m->rc6, m->rn6, rec->n_ring6, m->lc6, m->ln6,
                             lig->n_ring6);
  return e;
}

static int active(double c) { return c > 1e-5 || c < -1e-5; } /* utility.h:26 */

/* root_scorer::combine (scorer.h:108-117, scorer.cpp:53-126): inactive terms
 * (|coef| <= 1e-5) are neither evaluated nor combined; their term slot is 0. */
double or_pscore(const pxd_score_params* sp, const pxd_mol* rec, const double* rx,
                 const pxd_mol* lig, const double* lx, double* terms) {
  double t[PXD_NUM_TERMS] = {0};
  ring_memo m;
  ring_memo_fill(&m, rec, rx, lig, lx);
  if (active(sp->coul_coef)) t[PXD_TERM_COULOMB] = ps_coulomb(rec, rx, lig, lx);
  if (active(sp->vdw_coef)) t[PXD_TERM_VDW] = or_pscore_vdw(sp, rec, rx, lig, lx);
  if (active(sp->ionic_coef)) t[PXD_TERM_IONIC] = ps_ionic(sp, rec, rx, lig, lx);
  if (active(sp->pication_coef)) t[PXD_TERM_PICATION] = ps_cation_pi(sp, rec, rx, lig, lx, &m);
  if (active(sp->f2f_coef))
    t[PXD_TERM_F2F] = ps_stack(0.0, sp->f2f_sub_coef, sp->f2f_inter, rec, lig, &m);
  if (active(sp->e2f_coef))
    t[PXD_TERM_E2F] = ps_stack(90.0, sp->e2f_sub_coef, sp->e2f_inter, rec, lig, &m);
  if (active(sp->hbond_coef)) t[PXD_TERM_HBOND] = ps_hbond(sp, rec, rx, lig, lx);
  if (active(sp->hydrophobic_coef)) t[PXD_TERM_HYDROPHOBIC] = ps_hydrophobic(sp, rec, rx, lig, lx);
  if (active(sp->tstrain_coef)) t[PXD_TERM_TSTRAIN] = ps_tstrain(lig, lx);
  free(m.rc5);
  const double coef[PXD_NUM_TERMS] = {sp->coul_coef,     sp->vdw_coef,   sp->ionic_coef,
                                      sp->pication_coef, sp->f2f_coef,   sp->e2f_coef,
                                      sp->hbond_coef,    sp->hydrophobic_coef, sp->tstrain_coef};
  double s = 0.0;
  for (int k = 0; k < PXD_NUM_TERMS; ++k)
    if (active(coef[k])) s += coef[k] * t[k];
  if (terms) memcpy(terms, t, sizeof(t));
  return s;
}

/* ======================================================================== */
/* Affinity score: generalized Born (OBC) + surface area (scorer.cpp:735-866) */

void or_gbsa_atom_params(int32_t n, const int32_t* an, int32_t n_bonds, const int32_t* bonds,
                         double* radius, double* screen) {
  static const double radii[] = PXD_GB_RADII;
  static const double screens[] = PXD_GB_SCREENS;
  for (int i = 0; i < n; ++i) radius[i] = -1.0;
  for (int b = 0; b < n_bonds; ++b) {  /* scorer.cpp:742-749 */
    int i = bonds[2 * b], j = bonds[2 * b + 1];
    if (an[i] == 1 && an[j] == 7)
      radius[i] = PXD_GB_RADIUS_HN;
    else if (an[i] == 7 && an[j] == 1)
      radius[j] = PXD_GB_RADIUS_HN;
  }
  for (int i = 0; i < n; ++i) {  /* scorer.cpp:751-757 */
    if (radius[i] < 0.0) radius[i] = an[i] <= PXD_GB_NUM_RADII ? radii[an[i]] : radii[0];
    screen[i] = an[i] <= PXD_GB_NUM_SCREENS ? screens[an[i]] : screens[0];
  }
}

double or_gbsa_energy(int32_t n, const double* x, const double* q, const double* radius,
                      const double* screen) {
  double* orad = malloc(sizeof(double) * (n > 0 ? n : 1));
  double* sum_i = calloc(n > 0 ? n : 1, sizeof(double));
  double* born = malloc(sizeof(double) * (n > 0 ? n : 1));
  for (int i = 0; i < n; ++i) orad[i] = radius[i] - PXD_GB_OFFSET;
  /* descreening sums (scorer.cpp:807-830): source i, target j */
  for (int i = 0; i < n; ++i) {
    double sr = orad[i] * screen[i], sr2 = sr * sr;
    for (int j = 0; j < n; ++j) {
      if (i == j) continue;
      double r12 = dist3(x + 3 * i, x + 3 * j);
      double U = r12 + sr;
      if (U < orad[j]) continue;
      double r12_inv = 1.0 / r12;
      double D = fabs(r12 - sr);
      double L = orad[j] > D ? orad[j] : D;
      double U_inv = 1.0 / U, L_inv = 1.0 / L;
      D = 0.25 * (r12 - sr2 * r12_inv) * (U_inv * U_inv - L_inv * L_inv);
      D += L_inv - U_inv + 0.5 * log(L * U_inv + EPS) * r12_inv;
      D *= 0.5;
      sum_i[j] += D;
    }
  }
  /* Born radii (scorer.cpp:832-839) */
  for (int i = 0; i < n; ++i) {
    double psi = sum_i[i] * orad[i], psi2 = psi * psi;
    double th = tanh(PXD_GB_ALPHA * psi - PXD_GB_BETA * psi2 + PXD_GB_GAMMA * psi * psi2);
    born[i] = 1.0 / (1.0 / orad[i] - th / radius[i]);
  }
  /* generalized-Born energy over all ordered pairs, self terms included
   * (scorer.cpp:841-858) */
  double egb = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double f;
      if (i == j) {
        f = born[i];
      } else {
        double dx = x[3 * i] - x[3 * j], dy = x[3 * i + 1] - x[3 * j + 1],
               dz = x[3 * i + 2] - x[3 * j + 2];
        double r2 = dx * dx + dy * dy + dz * dz;  /* get_distance_square_3d */
        double bp = born[i] * born[j];
        f = sqrt(r2 + bp * exp(-0.25 * r2 / bp) + EPS);
      }
      double L = 0.5 * or_coulomb_factor *
                 (1.0 / PXD_GB_SOLUTE_DIELECTRIC - exp(-PXD_GB_KAPPA * f) / PXD_GB_SOLVENT_DIELECTRIC);
      egb -= q[i] * q[j] / f * L;
    }
  /* surface term (scorer.cpp:859-864) */
  double esa = 0.0;
  for (int i = 0; i < n; ++i) {
    double a = radius[i] + PXD_GB_PROBE_RADIUS;
    esa += a * a * pow(radius[i] / born[i], 6.0);
  }
  free(orad);
  free(sum_i);
  free(born);
  return egb + esa * PXD_GB_SURFACE_ENERGY;
}

double or_bscore(const pxd_bscore_params* bp, const pxd_mol* rec, const double* rx,
                 const double* rec_radius, const double* rec_screen, const pxd_mol* lig,
                 const double* lx, const double* lig_radius, const double* lig_screen,
                 double* terms) {
  double t[PXD_NUM_BSCORE_TERMS] = {0};
  double s = or_pscore(&bp->leaves, rec, rx, lig, lx, t);
  if (active(bp->gbsa_coef)) {
    const int nr = rec->n_atoms, nl = lig->n_atoms, nc = nr + nl;
    double* cx = malloc(sizeof(double) * 3 * nc);
    double* cq = malloc(sizeof(double) * nc);
    double* crad = malloc(sizeof(double) * nc);
    double* cscr = malloc(sizeof(double) * nc);
    /* the complex: ligand atoms first, then the receptor's (scorer.cpp:777-784) */
    memcpy(cx, lx, sizeof(double) * 3 * nl);
    memcpy(cx + 3 * nl, rx, sizeof(double) * 3 * nr);
    memcpy(cq, lig->charge, sizeof(double) * nl);
    memcpy(cq + nl, rec->charge, sizeof(double) * nr);
    memcpy(crad, lig_radius, sizeof(double) * nl);
    memcpy(crad + nl, rec_radius, sizeof(double) * nr);
    memcpy(cscr, lig_screen, sizeof(double) * nl);
    memcpy(cscr + nl, rec_screen, sizeof(double) * nr);
    double ep = or_gbsa_energy(nr, rx, rec->charge, rec_radius, rec_screen);
    double el = or_gbsa_energy(nl, lx, lig->charge, lig_radius, lig_screen);
    double ec = or_gbsa_energy(nc, cx, cq, crad, cscr);
    free(cx);
    free(cq);
    free(crad);
    free(cscr);
    t[PXD_NUM_TERMS] = ep;
    t[PXD_NUM_TERMS + 1] = el;
    t[PXD_NUM_TERMS + 2] = ec;
    s += bp->gbsa_coef * (ec - el - ep);
  }
  if (terms) memcpy(terms, t, sizeof(t));
  return s;
}

/* ======================================================================== */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11 "Parallel random numbers:
 * as easy as 1, 2, 3"), the generator shared with the GPU search kernels.   */

void or_philox4x32(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

/* ======================================================================== */
/* top-k clustering (writer.cpp:99-177), indices into the candidate list     */

static double rmsd_ub(const double* a, const double* b, int n) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) {
    double v[3];
    sub3(a + 3 * i, b + 3 * i, v);
    s += sqsum3(v);
  }
  return sqrt(s / n);
}

int32_t or_pose_cluster(int32_t n_cand, int32_t n, const double* xyz, const double* ps,
                        const int32_t* nevals, int32_t num_modes, double min_rmsd,
                        int32_t* out) {
  int cnt = 0;
  for (int i = 0; i < n_cand; ++i) {
    if (nevals[i] == 0) continue;
    if (cnt == 0) {
      out[cnt++] = i;
      continue;
    }
    int best = 0;
    double r = 0.0;
    for (int k = 0; k < cnt; ++k) {
      double v = rmsd_ub(xyz + 3 * (size_t)n * out[k], xyz + 3 * (size_t)n * i, n);
      if (k == 0 || v < r) {
        best = k;
        r = v;
      }
    }
    int updated = 0;
    if (r < min_rmsd) {
      if (ps[i] < ps[out[best]]) {
        out[best] = i;
        updated = 1;
      }
    } else if (cnt < num_modes) {
      out[cnt++] = i;
      updated = 1;
    } else if (ps[i] < ps[out[cnt - 1]]) {
      out[cnt - 1] = i;
      updated = 1;
    }
    if (updated) { /* stable insertion sort by pscore (std::sort on a short list) */
      for (int a = 1; a < cnt; ++a) {
        int v = out[a], b = a - 1;
        while (b >= 0 && ps[out[b]] > ps[v]) {
          out[b + 1] = out[b];
          --b;
        }
        out[b + 1] = v;
      }
    }
  }
  return cnt;
}

/* ======================================================================== */
/* Monte-Carlo walker (searcher.cpp:46-223) with the Philox draws of the GPU */
/* search: key (seed_lo, seed_hi + walker), counter (draw, ligand, 0, 0).    */

typedef struct mc_rng {
  uint32_t key[2];
  uint32_t ligand, ctr;
} mc_rng;

static void rng_block(mc_rng* r, uint32_t w[4]) {
  uint32_t c[4] = {r->ctr++, r->ligand, 0u, 0u};
  or_philox4x32(c, r->key, w);
}

static double rng_u01(mc_rng* r) {
  uint32_t w[4];
  rng_block(r, w);
  return (double)w[0] * 2.3283064365386963e-10; /* 2^-32 */
}

static double rng_uniform(mc_rng* r, double a, double b) { return a + (b - a) * rng_u01(r); }

static int rng_uniform_int(mc_rng* r, int a, int b) {
  int v = a + (int)floor(rng_u01(r) * (double)(b - a + 1));
  return v > b ? b : v;
}

static void rng_in_sphere(mc_rng* r, double v[3]) { /* random.cpp:42-49 */
  for (;;) {
    uint32_t w[4];
    rng_block(r, w);
    for (int d = 0; d < 3; ++d) v[d] = (double)w[d] * 4.656612873077393e-10 - 1.0;
    if (sqsum3(v) < 1.0) return;
  }
}

static void rng_quaternion(mc_rng* r, double q[4]) { /* random.cpp:29-40 */
  uint32_t w[4];
  rng_block(r, w);
  double u1 = (double)w[0] * 2.3283064365386963e-10;
  double u2 = (double)w[1] * 2.3283064365386963e-10 * 6.283185307179586;
  double u3 = (double)w[2] * 2.3283064365386963e-10 * 6.283185307179586;
  double s1 = sqrt(1.0 - u1), s2 = sqrt(u1);
  q[0] = s1 * sin(u2);
  q[1] = s1 * cos(u2);
  q[2] = s2 * sin(u3);
  q[3] = s2 * cos(u3);
}

static double rng_normal(mc_rng* r) { /* Box-Muller */
  uint32_t w[4];
  rng_block(r, w);
  double u1 = ((double)w[0] + 1.0) * 2.3283064365386963e-10;
  double u2 = (double)w[1] * 2.3283064365386963e-10;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

static int all_in_box(const double* x, int n, const pxd_box* b) { /* grid.h:42-46 */
  for (int a = 0; a < n; ++a)
    for (int d = 0; d < 3; ++d)
      if (!(x[3 * a + d] > b->init_xyz[d] && x[3 * a + d] < b->oppo_xyz[d])) return 0;
  return 1;
}

static void centroid3(const double* x, int n, double c[3]) {
  c[0] = c[1] = c[2] = 0.0;
  for (int a = 0; a < n; ++a) axpy3(c, 1.0, x + 3 * a);
  scale3(c, 1.0 / n);
}

static int coords_sane(const double* x, int n) { /* searcher.cpp:89-96 */
  for (int k = 0; k < 3 * n; ++k)
    if (!(x[k] <= 1e8 && x[k] >= -1e8)) return 0;
  return 1;
}

static double mc_pscore(const or_system* s, const pxd_score_params* sp, const double* lx,
                        const double* tors) {
  int nr = s->rec->n_atoms;
  double* rx = (double*)malloc(sizeof(double) * 3 * nr);
  if (s->leaves && s->leaves->n_leaves)
    or_receptor_apply(s->leaves, nr, s->rec_xyz, tors, rx); /* searcher.cpp:119 */
  else
    memcpy(rx, s->rec_xyz, sizeof(double) * 3 * nr);
  double v = or_pscore(sp, s->rec, rx, s->lig, lx, NULL);
  free(rx);
  return v;
}

/* randomize (searcher.cpp:46-87) from the input conformation xyz0.  The
 * torsions are drawn only when xyz0 is the ligand's conformer 0
 * (searcher.cpp:62-66); a walker's own pre-generated conformer keeps its. */
static void mc_randomize(const or_system* s, const pxd_frag_tree* t, const pxd_score_params* sp,
                         const pxd_mc_params* prm, const pxd_box* box, mc_rng* r,
                         const double* xyz0, int own_conformer, double* out) {
  int n = s->lig->n_atoms;
  double tors[PXD_MAX_FRAGMENTS] = {0};
  for (int tries = 1;; ++tries) {
    double cen[3], q[4], rot[9];
    for (int d = 0; d < 3; ++d) cen[d] = rng_uniform(r, box->init_xyz[d], box->oppo_xyz[d]);
    rng_quaternion(r, q);
    for (int k = 0; !own_conformer && k < t->n_torsions; ++k)
      tors[k] = rng_uniform(r, -3.141592653589793, 3.141592653589793);
    or_rotmat_quat(q, rot);
    or_ligand_apply(t, n, xyz0, cen, rot, tors, out);
    if (tries < prm->max_tries && !all_in_box(out, n, box)) continue;
    double cc[3];
    centroid3(out, n, cc);
    /* a centroid outside the box is retried; bounded at 4x max_tries */
    if (!all_in_box(cc, 1, box) && tries < 4 * prm->max_tries) continue;
    if (tries >= prm->max_tries) break;
    if (or_pscore_vdw(sp, s->rec, s->rec_xyz, s->lig, out) < 100.0) break;
  }
}

/* mutate_inplace (searcher.cpp:189-223), and the Gaussian variant
 * (mutation 1: translation and rotation-vector N(0, sigma^2), one torsion). */
static int mc_mutate(const pxd_frag_tree* t, const pxd_mc_params* prm, mc_rng* r, int n,
                     double* x) {
  double cen[3], rot[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, tors[PXD_MAX_FRAGMENTS] = {0};
  int which = -1;
  centroid3(x, n, cen);
  if (prm->mutation == 0) {
    which = rng_uniform_int(r, 0, t->n_torsions + 1);
    if (which == 0) {
      double dl[3];
      rng_in_sphere(r, dl);
      axpy3(cen, prm->amplitude, dl);
    } else if (which == 1) {
      double gr = 0.0;
      for (int a = 0; a < n; ++a) {
        double v[3];
        sub3(x + 3 * a, cen, v);
        gr += sqsum3(v);
      }
      gr = sqrt(gr / n);
      if (gr < 1e-6) return which; /* searcher.cpp:207 */
      double dl[3];
      rng_in_sphere(r, dl);
      scale3(dl, prm->amplitude / gr);
      double ang = norm3(dl);
      scale3(dl, 1.0 / ang);
      or_rotmat_axis(dl, ang, rot);
    } else {
      tors[which - 2] = rng_uniform(r, -3.141592653589793, 3.141592653589793);
    }
  } else {
    for (int d = 0; d < 3; ++d) cen[d] += rng_normal(r) * prm->trans_sigma;
    double v[3];
    for (int d = 0; d < 3; ++d) v[d] = rng_normal(r) * prm->rot_sigma;
    double ang = norm3(v);
    if (ang > 1e-12) {
      scale3(v, 1.0 / ang);
      or_rotmat_axis(v, ang, rot);
    }
    if (t->n_torsions > 0) {
      int k = rng_uniform_int(r, 0, t->n_torsions - 1);
      tors[k] = rng_uniform(r, -3.141592653589793, 3.141592653589793);
      which = k + 2;
    }
  }
  double* y = (double*)malloc(sizeof(double) * 3 * n);
  or_ligand_apply(t, n, x, cen, rot, tors, y);
  memcpy(x, y, sizeof(double) * 3 * n);
  free(y);
  return which;
}

int32_t or_mc_mutate(const pxd_frag_tree* t, const pxd_mc_params* prm, int32_t walker,
                     int32_t ligand, uint32_t* ctr, int32_t n, double* xyz) {
  mc_rng r = {{(uint32_t)prm->seed, (uint32_t)(prm->seed >> 32) + (uint32_t)walker},
              (uint32_t)ligand, *ctr};
  int which = mc_mutate(t, prm, &r, n, xyz);
  *ctr = r.ctr;
  return which;
}

void or_mc_draws(int32_t kind, uint64_t seed, int32_t walker, int32_t ligand, int32_t count,
                 double* out) {
  mc_rng r = {{(uint32_t)seed, (uint32_t)(seed >> 32) + (uint32_t)walker}, (uint32_t)ligand, 0u};
  for (int i = 0; i < count; ++i) {
    switch (kind) {
      case 0: out[i] = rng_u01(&r); break;
      case 1: rng_in_sphere(&r, out + 3 * i); break;
      case 2: rng_quaternion(&r, out + 4 * i); break;
      case 3: out[i] = rng_normal(&r); break;
      default: out[i] = (double)rng_uniform_int(&r, 0, kind - 4); break;
    }
  }
}

/* monte_carlo_searcher::mutate_and_optimize (searcher.cpp:98-187) for one
 * walker.  Returns the number of Monte-Carlo steps taken. */
int32_t or_mc_walker(const or_system* s, const pxd_frag_tree* t, const pxd_score_params* sp,
                     const pxd_mc_params* prm, const pxd_box* box, int32_t walker,
                     int32_t ligand, const double* xyz0, int32_t own_conformer,
                     double* emin_io, double* best_xyz, double* best_tors, double* best_e,
                     double* best_ps, int32_t* nevals, int32_t* n_pruned) {
  const int n = s->lig->n_atoms, n3 = 3 * n;
  const int nt = s->leaves ? s->leaves->n_leaves : 0;
  mc_rng r = {{(uint32_t)prm->seed, (uint32_t)(prm->seed >> 32) + (uint32_t)walker},
              (uint32_t)ligand, 0u};
  double* cur = (double*)malloc(sizeof(double) * 3 * n);
  double* cdd = (double*)malloc(sizeof(double) * 3 * n);
  double ct[PXD_MAX_LEAVES + 1] = {0}, dt[PXD_MAX_LEAVES + 1];
  if (walker == 0)
    memcpy(cur, xyz0, sizeof(double) * n3);
  else
    mc_randomize(s, t, sp, prm, box, &r, xyz0, own_conformer, cur);
  if (n_pruned) *n_pruned = 0;
  double* pg = (double*)malloc(sizeof(double) * 3 * n);
  double ptg[PXD_MAX_LEAVES + 1];
  double e, cur_ps;
  int32_t ne = 0;
  or_lbfgs(s, &prm->full, cur, ct, &e, &ne);
  cur_ps = mc_pscore(s, sp, cur, ct);
  *nevals = ne;
  *best_e = e;
  *best_ps = cur_ps;
  memcpy(best_xyz, cur, sizeof(double) * n3);
  if (nt) memcpy(best_tors, ct, sizeof(double) * nt);
  int steps = 0;
  if (walker != 0) {
    if (all_in_box(cur, n, box)) {
      for (int step = 0; step < prm->max_nsteps; ++step, ++steps) {
        memcpy(cdd, cur, sizeof(double) * n3);
        memcpy(dt, ct, sizeof(dt));
        mc_mutate(t, prm, &r, n, cdd);
        /* optional prune on the ligand's lowest pre-optimisation energy
         * (searcher.cpp:141-157; the minimum starts at 1e8, reader.cpp:51-53) */
        if (prm->mmenergy_threshold > 0.0 && emin_io) {
          double e0 = or_system_energy(s, dt, cdd, ptg, pg);
          if (e0 < *emin_io) *emin_io = e0;
          if (e0 > *emin_io + prm->mmenergy_threshold) {
            if (n_pruned) ++*n_pruned;
            continue;
          }
        }
        double ec, pc;
        or_lbfgs(s, &prm->relax, cdd, dt, &ec, &ne);
        *nevals += ne;
        pc = mc_pscore(s, sp, cdd, dt);
        /* Metropolis on pscore (searcher.cpp:164-165, searcher.h:53-56) */
        int acc = step == 0 || pc < cur_ps ||
                  rng_u01(&r) < exp((cur_ps - pc) / prm->temperature);
        if (!(acc && all_in_box(cdd, n, box))) continue;
        memcpy(cur, cdd, sizeof(double) * n3);
        memcpy(ct, dt, sizeof(dt));
        cur_ps = pc;
        if (cur_ps < *best_ps) { /* refine (searcher.cpp:167-174) */
          or_lbfgs(s, &prm->full, cur, ct, &ec, &ne);
          *nevals += ne;
          cur_ps = mc_pscore(s, sp, cur, ct);
          if (cur_ps < *best_ps) {
            *best_ps = cur_ps;
            *best_e = ec;
            memcpy(best_xyz, cur, sizeof(double) * n3);
            if (nt) memcpy(best_tors, ct, sizeof(double) * nt);
          }
        }
      }
    } else if (!coords_sane(cur, n)) {
      *nevals = 0; /* searcher.cpp:178-181 */
    }
  }
  free(cur);
  free(cdd);
  free(pg);
  return steps;
}
