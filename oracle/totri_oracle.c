/*
 * CPU ORACLE for the volumetric transient path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, operation for operation, the arithmetic of totrilib's
 * volume renderer, backprojections and RBF splat.  It is the checker for the
 * HIP kernels in totrilib_amd/csrc: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it.  The product path never calls it.
 *
 * Flavours (argument `mode`):
 *   0  FP32, same operations in the same order as the reference source, no
 *      contraction (compiled with -ffp-contract=off), sequential accumulation.
 *   1  FP64 weights and sums, with every discrete decision (d2 > 0, floor,
 *      ceil, range checks) taken from the FP32 flavour ("pinned").
 *   2  pure FP64, decisions taken in FP64 too (informational).
 * All outputs are written as double.
 *
 * Conventions fixed by the reference that this restatement keeps:
 *   powf(x, 2)  -> x*x        (IEEE product; parity with CUDA's powf unpinned)
 *   norm3df     -> sqrtf((dx*dx + dy*dy) + dz*dz)   (parity with CUDA unpinned)
 *   sqrtf, division: IEEE correctly rounded.
 *
 * Function-by-function citations point into the reference repository
 * (unlikelymaths/totrilib).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

int oracle_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void oracle_set_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

static inline float sq32(float x) { return x * x; }

/* ------------------------------------------------------------------------ */
/* R1: confocal volume render, src/render/volume_cuda.cu:11-66               */
/* ------------------------------------------------------------------------ */

/* FP32 sample decision and weights of one (scan point, bin, column) triple.
 * Returns 1 when the sample is active (d2 > 0 and 0 <= zi, zi+1 < Z),
 * following volume_cuda.cu:23,35,38,40-51,57. */
typedef struct {
  int active;
  int zi;
  float alpha;
  float depth;
  float f2;
  float d2;
} r1_sample32;

static inline float r1_tc32(int ti, float w, float off) {
  /* volume_cuda.cu:23  tc = ti * bin_width + bin_offset */
  float a = (float)ti * w;
  return a + off;
}

static inline float r1_x32(int xi, float o, float sz, int n) {
  /* volume_cuda.cu:35  origin + (xi + 0.5) * size / n  (mul, then div) */
  float a = ((float)xi + 0.5f) * sz;
  a = a / (float)n;
  return o + a;
}

static inline r1_sample32 r1_sample(float tc, float wx, float wy, float wz,
                                    float x, float y, float oz, float s, int Z) {
  r1_sample32 r;
  r.active = 0;
  r.zi = 0;
  r.alpha = 0.f;
  r.depth = 0.f;
  r.f2 = 0.f;
  /* volume_cuda.cu:40  depth^2 = tc^2 - (wx-x)^2 - (wy-y)^2, left to right */
  float t2 = sq32(tc);
  float dx = wx - x;
  float dy = wy - y;
  float d2 = t2 - sq32(dx);
  d2 = d2 - sq32(dy);
  r.d2 = d2;
  if (!(d2 > 0.f)) return r;
  float depth = sqrtf(d2);                 /* :42 */
  float z = wz + depth;                    /* :44 */
  float zh = (z - oz) * s;                 /* :46 */
  zh = zh - 0.5f;
  int zi = (int)floorf(zh);                /* :47 */
  if (!(zi >= 0 && zi + 1 < Z)) return r;  /* :49 */
  r.active = 1;
  r.zi = zi;
  r.alpha = zh - (float)zi;                /* :51 */
  r.depth = depth;
  /* :57  f2 = powf(-depth / powf(tc, 2), 4) */
  float q = -depth / t2;
  float q2 = q * q;
  r.f2 = q2 * q2;
  return r;
}

/* FP64 weights for a sample whose decisions were pinned in FP32. */
static inline void r1_weights64(int ti, float w, float off, float wx, float wy,
                                float wz, int xi, int yi, float ox, float oy,
                                float oz, float szx, float szy, float szz,
                                int X, int Y, int Z, int zi_pinned,
                                double* alpha, double* f2, double* zh_out,
                                double* d2_out) {
  double tc = (double)ti * (double)w + (double)off;
  double x = (double)ox + ((double)xi + 0.5) * (double)szx / (double)X;
  double y = (double)oy + ((double)yi + 0.5) * (double)szy / (double)Y;
  double s = (double)Z / (double)szz;
  double d2 = tc * tc - ((double)wx - x) * ((double)wx - x) -
              ((double)wy - y) * ((double)wy - y);
  double depth = d2 > 0 ? sqrt(d2) : 0.0;
  double zh = ((double)wz + depth - (double)oz) * s - 0.5;
  *alpha = zh - (double)zi_pinned;
  double q = depth / (tc * tc);
  *f2 = q * q * q * q;
  if (zh_out) *zh_out = zh;
  if (d2_out) *d2_out = d2;
}

/* transient[i][t] for i in [i0, i1).  volume [X][Y][Z], mp [3][N],
 * out [i1 - i0][T] (double).  With cols != NULL only the ncols listed (xi, yi)
 * columns are summed: the exact render of a volume that is zero elsewhere
 * (a zero column adds exact zeros in every mode). */
void oracle_r1_forward(const float* vol, int X, int Y, int Z,
                       const float* mp, int N, int T,
                       const float* origin, const float* size,
                       float w, float off, int mode,
                       double* out, int i0, int i1, const int* cols, int ncols) {
  const float ox = origin[0], oy = origin[1], oz = origin[2];
  const float szx = size[0], szy = size[1], szz = size[2];
  const float s = (float)Z / szz; /* :30 */
  const double s64 = (double)Z / (double)szz;
  if (i0 < 0) i0 = 0;
  if (i1 > N) i1 = N;
#pragma omp parallel for collapse(2) schedule(dynamic, 16)
  for (int i = i0; i < i1; ++i) {
    for (int ti = 0; ti < T; ++ti) {
      const float wx = mp[0 * N + i], wy = mp[1 * N + i], wz = mp[2 * N + i];
      const float tc = r1_tc32(ti, w, off);
      float v32 = 0.f;
      double v64 = 0.0;
      const double tc64 = (double)ti * (double)w + (double)off;
      const int nc = cols ? ncols : X * Y;
      for (int c = 0; c < nc; ++c) {
        const int xi = cols ? cols[2 * c] : c / Y;
        const int yi = cols ? cols[2 * c + 1] : c % Y;
        const float x = r1_x32(xi, ox, szx, X);
        const double x64 = (double)ox + ((double)xi + 0.5) * (double)szx / (double)X;
        {
          const float y = r1_x32(yi, oy, szy, Y);
          const float* col = vol + ((size_t)xi * Y + yi) * Z;
          if (mode == 2) {
            const double y64 = (double)oy + ((double)yi + 0.5) * (double)szy / (double)Y;
            double d2 = tc64 * tc64 - ((double)wx - x64) * ((double)wx - x64) -
                        ((double)wy - y64) * ((double)wy - y64);
            if (d2 > 0) {
              double depth = sqrt(d2);
              double zh = ((double)wz + depth - (double)oz) * s64 - 0.5;
              double fl = floor(zh);
              if (fl >= 0 && fl + 1 < Z) {
                int zi = (int)fl;
                double a = zh - fl;
                double f1 = (1 - a) * col[zi] + a * col[zi + 1];
                double q = depth / (tc64 * tc64);
                v64 += f1 * (q * q * q * q);
              }
            }
            continue;
          }
          r1_sample32 smp = r1_sample(tc, wx, wy, wz, x, y, oz, s, Z);
          if (!smp.active) continue;
          if (mode == 0) {
            /* :53-59 */
            const float v0 = col[smp.zi], v1 = col[smp.zi + 1];
            float f1 = (1.f - smp.alpha) * v0;
            f1 = f1 + smp.alpha * v1;
            v32 = v32 + f1 * smp.f2;
          } else {
            double a, f2;
            r1_weights64(ti, w, off, wx, wy, wz, xi, yi, ox, oy, oz, szx, szy,
                         szz, X, Y, Z, smp.zi, &a, &f2, NULL, NULL);
            double f1 = (1 - a) * col[smp.zi] + a * col[smp.zi + 1];
            v64 += f1 * f2;
          }
        }
      }
      out[(size_t)(i - i0) * T + ti] = (mode == 0) ? (double)v32 : v64;
    }
  }
}

/* Per-sample decision dump for (i, t, xi, yi) quadruples, FP32 flavour.
 * zi_out = -1000000 for inactive samples with d2 <= 0, otherwise floor(z_hat)
 * (also when out of range).  Used to check the kernels' bin decisions
 * bit for bit. */
void oracle_r1_decisions(const int* idx, int n, const float* mp, int N,
                         int X, int Y, int Z, const float* origin,
                         const float* size, float w, float off,
                         int* zi_out, float* alpha_out, float* d2_out) {
  const float s = (float)Z / size[2];
  for (int k = 0; k < n; ++k) {
    const int i = idx[4 * k], ti = idx[4 * k + 1], xi = idx[4 * k + 2],
              yi = idx[4 * k + 3];
    const float wx = mp[i], wy = mp[N + i], wz = mp[2 * N + i];
    const float tc = r1_tc32(ti, w, off);
    const float x = r1_x32(xi, origin[0], size[0], X);
    const float y = r1_x32(yi, origin[1], size[1], Y);
    float t2 = sq32(tc);
    float dx = wx - x, dy = wy - y;
    float d2 = t2 - sq32(dx);
    d2 = d2 - sq32(dy);
    d2_out[k] = d2;
    if (!(d2 > 0.f)) {
      zi_out[k] = -1000000;
      alpha_out[k] = 0.f;
      continue;
    }
    float depth = sqrtf(d2);
    float z = wz + depth;
    float zh = (z - origin[2]) * s;
    zh = zh - 0.5f;
    float fl = floorf(zh);
    zi_out[k] = (int)fl;
    alpha_out[k] = zh - fl;
  }
}

/* ------------------------------------------------------------------------ */
/* A1 (build): exact transpose of R1.  vol[c][z] = sum over samples (i,t,c)  */
/* of (1-alpha) f2 G[i][t] into zi and alpha f2 G[i][t] into zi+1.           */
/* Columns listed in `cols` (pairs xi, yi); all when cols == NULL.            */
/* out is [X][Y][Z] (double); only listed columns are written.                */
/* ------------------------------------------------------------------------ */
void oracle_a1_transpose(const double* G, const float* mp, int N, int T,
                         int X, int Y, int Z, const float* origin,
                         const float* size, float w, float off, int mode,
                         const int* cols, int ncols, double* out) {
  const float ox = origin[0], oy = origin[1], oz = origin[2];
  const float szx = size[0], szy = size[1], szz = size[2];
  const float s = (float)Z / szz;
  const double s64 = (double)Z / (double)szz;
  const int nc = cols ? ncols : X * Y;
#pragma omp parallel for schedule(dynamic, 1)
  for (int k = 0; k < nc; ++k) {
    const int xi = cols ? cols[2 * k] : k / Y;
    const int yi = cols ? cols[2 * k + 1] : k % Y;
    double* o = out + ((size_t)xi * Y + yi) * Z;
    float* o32 = (float*)malloc(sizeof(float) * (size_t)Z);
    for (int z = 0; z < Z; ++z) { o[z] = 0.0; o32[z] = 0.f; }
    const float x = r1_x32(xi, ox, szx, X);
    const float y = r1_x32(yi, oy, szy, Y);
    const double x64 = (double)ox + ((double)xi + 0.5) * (double)szx / (double)X;
    const double y64 = (double)oy + ((double)yi + 0.5) * (double)szy / (double)Y;
    for (int i = 0; i < N; ++i) {
      const float wx = mp[i], wy = mp[N + i], wz = mp[2 * N + i];
      for (int ti = 0; ti < T; ++ti) {
        const double g = G[(size_t)i * T + ti];
        if (mode == 2) {
          const double tc64 = (double)ti * (double)w + (double)off;
          double d2 = tc64 * tc64 - ((double)wx - x64) * ((double)wx - x64) -
                      ((double)wy - y64) * ((double)wy - y64);
          if (!(d2 > 0)) continue;
          double depth = sqrt(d2);
          double zh = ((double)wz + depth - (double)oz) * s64 - 0.5;
          double fl = floor(zh);
          if (!(fl >= 0 && fl + 1 < Z)) continue;
          int zi = (int)fl;
          double a = zh - fl;
          double q = depth / (tc64 * tc64);
          double f2 = q * q * q * q;
          o[zi] += (1 - a) * f2 * g;
          o[zi + 1] += a * f2 * g;
          continue;
        }
        const float tc = r1_tc32(ti, w, off);
        r1_sample32 smp = r1_sample(tc, wx, wy, wz, x, y, oz, s, Z);
        if (!smp.active) continue;
        if (mode == 0) {
          const float g32 = (float)g;
          const float wg = smp.f2 * g32;
          o32[smp.zi] += (1.f - smp.alpha) * wg;
          o32[smp.zi + 1] += smp.alpha * wg;
        } else {
          double a, f2;
          r1_weights64(ti, w, off, wx, wy, wz, xi, yi, ox, oy, oz, szx, szy,
                       szz, X, Y, Z, smp.zi, &a, &f2, NULL, NULL);
          o[smp.zi] += (1 - a) * f2 * g;
          o[smp.zi + 1] += a * f2 * g;
        }
      }
    }
    if (mode == 0)
      for (int z = 0; z < Z; ++z) o[z] = (double)o32[z];
    free(o32);
  }
}

/* ------------------------------------------------------------------------ */
/* A1 (reference window form): volume_cuda.cu:95-160, restated as written.  */
/* FP32 only (mode 0) or FP64 weights with pinned windows (mode 1).          */
/* ------------------------------------------------------------------------ */
void oracle_a1_reference(const double* G, const float* mp, int N, int T,
                         int X, int Y, int Z, const float* origin,
                         const float* size, float w, float off, int mode,
                         const int* cols, int ncols, double* out) {
  const float ox = origin[0], oy = origin[1], oz = origin[2];
  const float szx = size[0], szy = size[1], szz = size[2];
  const float s = (float)Z / szz;   /* :112 */
  const float dz = szz / (float)Z;  /* :114 */
  const int nc = cols ? ncols : X * Y;
#pragma omp parallel for schedule(dynamic, 1)
  for (int k = 0; k < nc; ++k) {
    const int xi = cols ? cols[2 * k] : k / Y;
    const int yi = cols ? cols[2 * k + 1] : k % Y;
    const float x = r1_x32(xi, ox, szx, X); /* :107 */
    const float y = r1_x32(yi, oy, szy, Y); /* :108 */
    for (int zi = 0; zi < Z; ++zi) {
      const float z = r1_x32(zi, oz, szz, Z); /* :109 */
      float v32 = 0.f;
      double v64 = 0.0;
      for (int i = 0; i < N; ++i) {
        const float wx = mp[i], wy = mp[N + i], wz = mp[2 * N + i];
        const float zz = zi > 0 ? dz : 0.f;                  /* :123 */
        float dps = sq32(x - wx);
        dps = dps + sq32(y - wy);                            /* :124 */
        const float ds = sqrtf(dps + sq32((z - zz) - wz));   /* :125 */
        const float dm = sqrtf(dps + sq32(z - wz));          /* :126 */
        const float de = sqrtf(dps + sq32((z + dz) - wz));   /* :127 */
        const float tfs = (ds - off) / w;                    /* :129-131 */
        const float tfm = (dm - off) / w;
        const float tfe = (de - off) / w;
        int ts = (int)ceilf(tfs), tm = (int)ceilf(tfm), te = (int)ceilf(tfe);
        ts = ts < 0 ? 0 : (ts > T ? T : ts);                 /* :132-134 */
        tm = tm < 0 ? 0 : (tm > T ? T : tm);
        te = te < 0 ? 0 : (te > T ? T : te);
        for (int ti = ts; ti < te; ++ti) {                   /* :135 */
          const float tc = r1_tc32(ti, w, off);              /* :137 */
          const float t2 = sq32(tc);
          const float d2 = t2 - dps;                         /* :139 */
          if (!(d2 > 0.f)) continue;
          const float depth = sqrtf(d2);
          const float zc = wz + depth;                       /* :143 */
          float zh = (zc - oz) * s;
          zh = zh - 0.5f;                                    /* :145 */
          const int zf = (int)floorf(zh);
          const float alpha = zh - (float)zf;                /* :148 */
          const double g = G[(size_t)i * T + ti];
          if (mode == 0) {
            const float df1 = (ti >= tm) ? (1.f - alpha) : alpha; /* :150 */
            float q = -depth / t2;
            float q2 = q * q;
            const float f2 = q2 * q2;                        /* :152 */
            v32 = v32 + (df1 * f2) * (float)g;               /* :154 */
          } else {
            double tc64 = (double)ti * (double)w + (double)off;
            double x64 = (double)ox + ((double)xi + 0.5) * (double)szx / (double)X;
            double y64 = (double)oy + ((double)yi + 0.5) * (double)szy / (double)Y;
            double d2d = tc64 * tc64 - (x64 - wx) * (x64 - wx) - (y64 - wy) * (y64 - wy);
            double dep = d2d > 0 ? sqrt(d2d) : 0.0;
            double zh64 = ((double)wz + dep - (double)oz) * ((double)Z / (double)szz) - 0.5;
            double a64 = zh64 - (double)zf;
            double df1 = (ti >= tm) ? (1 - a64) : a64;
            double q = dep / (tc64 * tc64);
            v64 += df1 * q * q * q * q * g;
          }
        }
      }
      out[((size_t)xi * Y + yi) * Z + zi] = (mode == 0) ? (double)v32 : v64;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* A6: gradient w.r.t. measurement points, volume_cuda.cu:188-265.           */
/* out [3][N] (double), FP32 flavour (0) or FP64 weights, pinned (1).        */
/* ------------------------------------------------------------------------ */
void oracle_a6(const float* vol, int X, int Y, int Z, const float* mp, int N,
               const double* G, int T, const float* origin, const float* size,
               float w, float off, int mode, double* out, int i0, int i1) {
  const float ox = origin[0], oy = origin[1], oz = origin[2];
  const float szx = size[0], szy = size[1], szz = size[2];
  const float s = (float)Z / szz;
  const double s64 = (double)Z / (double)szz;
  if (i0 < 0) i0 = 0;
  if (i1 > N) i1 = N;
#pragma omp parallel for schedule(dynamic, 1)
  for (int i = i0; i < i1; ++i) {
    const float wx = mp[i], wy = mp[N + i], wz = mp[2 * N + i];
    double gx = 0, gy = 0, gz = 0;
    float gx32 = 0, gy32 = 0, gz32 = 0;
    for (int ti = 0; ti < T; ++ti) {
      const double dl = G[(size_t)i * T + ti];
      const float tc = r1_tc32(ti, w, off);
      float ax = 0, ay = 0, az = 0;
      double bx = 0, by = 0, bz = 0;
      for (int xi = 0; xi < X; ++xi) {
        const float x = r1_x32(xi, ox, szx, X);
        for (int yi = 0; yi < Y; ++yi) {
          const float y = r1_x32(yi, oy, szy, Y);
          r1_sample32 smp = r1_sample(tc, wx, wy, wz, x, y, oz, s, Z);
          if (!smp.active) continue;
          const float* col = vol + ((size_t)xi * Y + yi) * Z;
          const float v0 = col[smp.zi], v1 = col[smp.zi + 1];
          if (mode == 0) {
            const float depth = smp.depth;
            const float dzdwx = -(wx - x) / depth;            /* :228 */
            const float dzdwy = -(wy - y) / depth;            /* :229 */
            const float dzdwz = 1.f;
            float f1 = (1.f - smp.alpha) * v0;
            f1 = f1 + smp.alpha * v1;                         /* :241 */
            const float dv = v1 - v0;
            const float df1x = (s * dzdwx) * dv;              /* :243 */
            const float df1y = (s * dzdwy) * dv;
            const float df1z = (s * dzdwz) * dv;
            /* :249 tmp = 4 * powf(-depth / powf(tc, 8/3), 3)
               = -4 depth^3 / tc^8, written with exact powers. */
            const float t2 = tc * tc;
            const float t8 = (t2 * t2) * (t2 * t2);
            const float tmp = -4.f * (depth * depth * depth) / t8;
            const float df2x = -dzdwx * tmp;                  /* :250 */
            const float df2y = -dzdwy * tmp;
            const float df2z = (1.f - dzdwz) * tmp;
            ax += f1 * df2x + df1x * smp.f2;                  /* :254 */
            ay += f1 * df2y + df1y * smp.f2;
            az += f1 * df2z + df1z * smp.f2;
          } else {
            double a, f2, zh, d2;
            r1_weights64(ti, w, off, wx, wy, wz, xi, yi, ox, oy, oz, szx, szy,
                         szz, X, Y, Z, smp.zi, &a, &f2, &zh, &d2);
            double tc64 = (double)ti * (double)w + (double)off;
            double x64 = (double)ox + ((double)xi + 0.5) * (double)szx / (double)X;
            double y64 = (double)oy + ((double)yi + 0.5) * (double)szy / (double)Y;
            double depth = sqrt(d2 > 0 ? d2 : 0);
            double dzdwx = -((double)wx - x64) / depth;
            double dzdwy = -((double)wy - y64) / depth;
            double f1 = (1 - a) * v0 + a * v1;
            double dv = (double)v1 - (double)v0;
            double t8 = pow(tc64, 8);
            double tmp = -4.0 * depth * depth * depth / t8;
            bx += f1 * (-dzdwx * tmp) + s64 * dzdwx * dv * f2;
            by += f1 * (-dzdwy * tmp) + s64 * dzdwy * dv * f2;
            bz += s64 * dv * f2;
          }
        }
      }
      if (mode == 0) {
        gx32 += (float)dl * ax;                               /* :261-263 */
        gy32 += (float)dl * ay;
        gz32 += (float)dl * az;
      } else {
        gx += dl * bx;
        gy += dl * by;
        gz += dl * bz;
      }
    }
    out[i] = mode == 0 ? gx32 : gx;
    out[N + i] = mode == 0 ? gy32 : gy;
    out[2 * N + i] = mode == 0 ? gz32 : gz;
  }
}

/* ------------------------------------------------------------------------ */
/* norm3df restatement: sqrtf((dx*dx + dy*dy) + dz*dz)                       */
/* ------------------------------------------------------------------------ */
static inline float dist32(float ax, float ay, float az, float bx, float by,
                           float bz) {
  const float dx = ax - bx, dy = ay - by, dz = az - bz;
  float s = dx * dx;
  s = s + dy * dy;
  s = s + dz * dz;
  return sqrtf(s);
}
static inline double dist64(double ax, double ay, double az, double bx,
                            double by, double bz) {
  const double dx = ax - bx, dy = ay - by, dz = az - bz;
  return sqrt(dx * dx + dy * dy + dz * dz);
}

/* ------------------------------------------------------------------------ */
/* R2: exhaustive volume render, volume_cuda.cu:297-377.                     */
/* volume [B][D][H][W], scan [B][3][S], laser [B][3][L], out [B][T][S][L].   */
/* Contributions are summed per output element in (voxel row-major, zi)     */
/* order.  mode 0: FP32 (weights and sums); mode 1: FP32 decisions and       */
/* weights, FP64 sums.                                                        */
/* ------------------------------------------------------------------------ */
static void r2_voxel_bins(float occupancy, float xf, float yf, float zf0,
                          float zf1, const float sp[3], const float lp[3],
                          float w, float off, int T, int* first, int* last,
                          float* wfirst, float* wmid, float* wlast, int* single) {
  float d0 = dist32(sp[0], sp[1], sp[2], xf, yf, zf0) +
             dist32(lp[0], lp[1], lp[2], xf, yf, zf0);  /* :339 */
  float d1 = dist32(sp[0], sp[1], sp[2], xf, yf, zf1) +
             dist32(lp[0], lp[1], lp[2], xf, yf, zf1);  /* :340 */
  if (d1 < d0) { float t = d0; d0 = d1; d1 = t; }       /* :341-345 */
  float bf0 = (d0 - off) / w;                            /* :346 */
  float bf1 = (d1 - off) / w;                            /* :347 */
  const float range = bf1 - bf0;                         /* :348 */
  int fb = (int)floorf(bf0), lb = (int)floorf(bf1);      /* :349-350 */
  *first = -1;
  *last = -2;
  if (fb < 0) {                                          /* :351-357 */
    if (lb < 0) return;
    bf0 = 0.f;
    fb = 0;
  }
  if (lb >= T) {                                         /* :358-364 */
    if (fb >= T) return;
    bf1 = (float)(T - 1);
    lb = T - 1;
  }
  *first = fb;
  *last = lb;
  if (fb == lb) {                                        /* :365-366 */
    *single = 1;
    *wfirst = occupancy;
    return;
  }
  *single = 0;
  const float frac = occupancy / range;                  /* :368 */
  *wfirst = (1.f - (bf0 - (float)fb)) * frac;            /* :369 */
  *wmid = frac;                                          /* :371 */
  *wlast = (bf1 - (float)lb) * frac;                     /* :373 */
}

static inline void r2_geometry(int xi, int yi, int zi, int D, int H, int W,
                               const float* vs, const float* ve, float* xf,
                               float* yf, float* zf0, float* zf1) {
  float a = ((float)xi + 0.5f) / (float)W;               /* :316-317 */
  *xf = vs[0] + a * (ve[0] - vs[0]);
  a = ((float)yi + 0.5f) / (float)H;                     /* :318-319 */
  *yf = vs[1] + a * (ve[1] - vs[1]);
  a = ((float)zi + 0.5f) / (float)D;                     /* :335-336 */
  const float zf = vs[2] + a * (ve[2] - vs[2]);
  const float h = (0.5f / (float)D) * (ve[2] - vs[2]);   /* :337-338 */
  *zf0 = zf - h;
  *zf1 = zf + h;
}


/* FP64 weights of R2's box spread with the FP32 decisions pinned: the bin
 * span [fb, lb], the single-bin case and the clipping of bf0 / bf1 at the
 * histogram ends are those of r2_voxel_bins. */
static void r2_weights64(double occ, int xi, int yi, int zi, int D, int H,
                         int W, const float* vs, const float* ve,
                         const float sp[3], const float lp[3], float w,
                         float off, int T, int fb, int lb, int single,
                         double* wf, double* wm, double* wl) {
  const double xf = (double)vs[0] + ((double)xi + 0.5) / W * ((double)ve[0] - vs[0]);
  const double yf = (double)vs[1] + ((double)yi + 0.5) / H * ((double)ve[1] - vs[1]);
  const double zf = (double)vs[2] + ((double)zi + 0.5) / D * ((double)ve[2] - vs[2]);
  const double h = 0.5 / D * ((double)ve[2] - vs[2]);
  double d0 = dist64(sp[0], sp[1], sp[2], xf, yf, zf - h) +
              dist64(lp[0], lp[1], lp[2], xf, yf, zf - h);
  double d1 = dist64(sp[0], sp[1], sp[2], xf, yf, zf + h) +
              dist64(lp[0], lp[1], lp[2], xf, yf, zf + h);
  if (d1 < d0) { double t = d0; d0 = d1; d1 = t; }
  double bf0 = (d0 - off) / w, bf1 = (d1 - off) / w;
  const double range = bf1 - bf0;
  if (single) {
    *wf = occ;
    return;
  }
  /* clipping pinned from FP32: a clipped end is recognised by its bin */
  const int clipped0 = (fb == 0) && (floor(bf0) < 0);
  const int clipped1 = (lb == T - 1) && (floor(bf1) >= T);
  if (clipped0) bf0 = 0.0;
  if (clipped1) bf1 = (double)(T - 1);
  const double frac = occ / range;
  *wf = (1.0 - (bf0 - fb)) * frac;
  *wm = frac;
  *wl = (bf1 - lb) * frac;
}

void oracle_r2_forward(const float* vol, int B, int D, int H, int W,
                       const float* scan, int S, const float* laser, int L,
                       int T, const float* vs, const float* ve, float w,
                       float off, int mode, double* out) {
  const size_t per_b = (size_t)T * S * L;
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
  for (int b = 0; b < B; ++b) {
    for (int s = 0; s < S; ++s) {
      float* acc32 = (float*)calloc((size_t)T * L, sizeof(float));
      double* acc64 = (double*)calloc((size_t)T * L, sizeof(double));
      for (int l = 0; l < L; ++l) {
        const float sp[3] = {scan[(size_t)b * 3 * S + s],
                             scan[(size_t)b * 3 * S + S + s],
                             scan[(size_t)b * 3 * S + 2 * S + s]};
        const float lp[3] = {laser[(size_t)b * 3 * L + l],
                             laser[(size_t)b * 3 * L + L + l],
                             laser[(size_t)b * 3 * L + 2 * L + l]};
        for (int v = 0; v < H * W; ++v) {
          const int xi = v % W, yi = v / W;
          for (int zi = 0; zi < D; ++zi) {
            const float occ = vol[(((size_t)b * D + zi) * H + yi) * W + xi];
            if (occ == 0.f) continue;                    /* :332-334 */
            float xf, yf, zf0, zf1;
            r2_geometry(xi, yi, zi, D, H, W, vs, ve, &xf, &yf, &zf0, &zf1);
            int fb, lb, single;
            float wf, wm, wl;
            r2_voxel_bins(occ, xf, yf, zf0, zf1, sp, lp, w, off, T, &fb, &lb,
                          &wf, &wm, &wl, &single);
            if (fb > lb) continue;
            if (mode == 2) {
              double df, dm, dl;
              r2_weights64(occ, xi, yi, zi, D, H, W, vs, ve, sp, lp, w, off,
                           T, fb, lb, single, &df, &dm, &dl);
              if (single) {
                acc64[(size_t)fb * L + l] += df;
              } else {
                acc64[(size_t)fb * L + l] += df;
                for (int bn = fb + 1; bn < lb; ++bn) acc64[(size_t)bn * L + l] += dm;
                acc64[(size_t)lb * L + l] += dl;
              }
            } else if (single) {
              acc32[(size_t)fb * L + l] += wf;
              acc64[(size_t)fb * L + l] += wf;
            } else {
              acc32[(size_t)fb * L + l] += wf;
              acc64[(size_t)fb * L + l] += wf;
              for (int bn = fb + 1; bn < lb; ++bn) {
                acc32[(size_t)bn * L + l] += wm;
                acc64[(size_t)bn * L + l] += wm;
              }
              acc32[(size_t)lb * L + l] += wl;
              acc64[(size_t)lb * L + l] += wl;
            }
          }
        }
      }
      for (int t = 0; t < T; ++t)
        for (int l = 0; l < L; ++l)
          out[(size_t)b * per_b + ((size_t)t * S + s) * L + l] =
              mode == 0 ? (double)acc32[(size_t)t * L + l] : acc64[(size_t)t * L + l];
      free(acc32);
      free(acc64);
    }
  }
}

/* A5 (build): true adjoint of R2 (the reference kernel body is empty,
 * volume_cuda.cu:414-431).  grad[b][zi][yi][xi] = sum over (s,l) of the R2
 * bin weights of a unit-occupancy voxel times G[b][bin][s][l].
 * FP32 decisions; weights in FP32 (mode 0) or recomputed in FP64 (mode 1). */
void oracle_r2_adjoint(const double* G, int B, int D, int H, int W,
                       const float* scan, int S, const float* laser, int L,
                       int T, const float* vs, const float* ve, float w,
                       float off, int mode, double* out) {
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
  for (int b = 0; b < B; ++b) {
    for (int v = 0; v < H * W; ++v) {
      const int xi = v % W, yi = v / W;
      for (int zi = 0; zi < D; ++zi) {
        float xf, yf, zf0, zf1;
        r2_geometry(xi, yi, zi, D, H, W, vs, ve, &xf, &yf, &zf0, &zf1);
        double acc = 0.0;
        float acc32 = 0.f;
        for (int s = 0; s < S; ++s) {
          const float sp[3] = {scan[(size_t)b * 3 * S + s],
                               scan[(size_t)b * 3 * S + S + s],
                               scan[(size_t)b * 3 * S + 2 * S + s]};
          for (int l = 0; l < L; ++l) {
            const float lp[3] = {laser[(size_t)b * 3 * L + l],
                                 laser[(size_t)b * 3 * L + L + l],
                                 laser[(size_t)b * 3 * L + 2 * L + l]};
            int fb, lb, single;
            float wf, wm, wl;
            r2_voxel_bins(1.f, xf, yf, zf0, zf1, sp, lp, w, off, T, &fb, &lb,
                          &wf, &wm, &wl, &single);
            if (fb > lb) continue;
            const double* g = G + (size_t)b * T * S * L;
#define GG(bn) g[((size_t)(bn) * S + s) * L + l]
            if (single) {
              acc += (double)wf * GG(fb);
              acc32 += wf * (float)GG(fb);
            } else {
              acc += (double)wf * GG(fb);
              acc32 += wf * (float)GG(fb);
              for (int bn = fb + 1; bn < lb; ++bn) {
                acc += (double)wm * GG(bn);
                acc32 += wm * (float)GG(bn);
              }
              acc += (double)wl * GG(lb);
              acc32 += wl * (float)GG(lb);
            }
#undef GG
          }
        }
        out[(((size_t)b * D + zi) * H + yi) * W + xi] = mode == 0 ? (double)acc32 : acc;
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* A2: confocal Velten backprojection, backprojection_cuda.cu:152-194.       */
/* transient [B][T][S], scan [B][S][3], origin [B][So][3] with So in {0,S}.  */
/* out [B][D][H][W].  mode 0: FP32 sums; 1: FP32 decisions, FP64 sums.       */
/* ------------------------------------------------------------------------ */
static inline float inclusive32(int i, int n, float a, float b) {
  /* backprojection_cuda.cu:168-170  start + (i / (n-1)) * (end - start) */
  float t = (float)i / (float)(n - 1);
  return a + t * (b - a);
}

void oracle_velten_confocal(const double* tr, int B, int T, int S,
                            const float* scan, const float* sorigin, int So,
                            int D, int H, int W, const float* vs,
                            const float* ve, float w, float off, int mode,
                            double* out) {
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
  for (int b = 0; b < B; ++b) {
    for (int z = 0; z < D; ++z) {
      for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
          const float qx = inclusive32(x, W, vs[0], ve[0]);
          const float qy = inclusive32(y, H, vs[1], ve[1]);
          const float qz = inclusive32(z, D, vs[2], ve[2]);
          float v32 = 0.f;
          double v64 = 0.0;
          for (int s = 0; s < S; ++s) {
            const float* p = scan + ((size_t)b * S + s) * 3;
            const float dsv = dist32(p[0], p[1], p[2], qx, qy, qz); /* :177 */
            float dso = 0.f;
            if (So > 0) {
              const float* o = sorigin + ((size_t)b * So + s) * 3;
              dso = dist32(p[0], p[1], p[2], o[0], o[1], o[2]);     /* :178-183 */
            }
            const float dist = 2.f * dso + 2.f * dsv;               /* :185 */
            float tf = (dist - off) / w;
            tf = tf + 0.5f;
            const int ti = (int)floorf(tf);                         /* :187 */
            if (ti >= 0 && ti < T) {
              const double g = tr[((size_t)b * T + ti) * S + s];
              v32 += (float)g;
              v64 += g;
            }
          }
          out[(((size_t)b * D + z) * H + y) * W + x] = mode == 0 ? (double)v32 : v64;
        }
      }
    }
  }
}

/* A3: exhaustive Velten, backprojection_cuda.cu:228-287.
 * transient [B][T][S][L], scan [B][S][3], laser [B][L][3]. */
void oracle_velten_exhaustive(const double* tr, int B, int T, int S, int L,
                              const float* scan, const float* laser,
                              const float* sorigin, int So,
                              const float* lorigin, int Lo, int D, int H,
                              int W, const float* vs, const float* ve,
                              float w, float off, int mode, double* out) {
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
  for (int b = 0; b < B; ++b) {
    for (int z = 0; z < D; ++z) {
      for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
          const float qx = inclusive32(x, W, vs[0], ve[0]);
          const float qy = inclusive32(y, H, vs[1], ve[1]);
          const float qz = inclusive32(z, D, vs[2], ve[2]);
          float v32 = 0.f;
          double v64 = 0.0;
          for (int s = 0; s < S; ++s) {
            const float* p = scan + ((size_t)b * S + s) * 3;
            const float dsv = dist32(p[0], p[1], p[2], qx, qy, qz);
            float dso = 0.f;
            if (So > 0) {
              const float* o = sorigin + ((size_t)b * So + s) * 3;
              dso = dist32(p[0], p[1], p[2], o[0], o[1], o[2]);
            }
            for (int l = 0; l < L; ++l) {
              const float* lp = laser + ((size_t)b * L + l) * 3;
              const float dlv = dist32(lp[0], lp[1], lp[2], qx, qy, qz);
              float dlo = 0.f;
              if (Lo > 0) {
                const float* o = lorigin + ((size_t)b * Lo + l) * 3;
                dlo = dist32(lp[0], lp[1], lp[2], o[0], o[1], o[2]);
              }
              float dist = dso + dsv;                               /* :276-277 */
              dist = dist + dlv;
              dist = dist + dlo;
              float tf = (dist - off) / w;
              tf = tf + 0.5f;
              const int ti = (int)floorf(tf);                       /* :279 */
              if (ti >= 0 && ti < T) {
                const double g = tr[(((size_t)b * T + ti) * S + s) * L + l];
                v32 += (float)g;
                v64 += g;
              }
            }
          }
          out[(((size_t)b * D + z) * H + y) * W + x] = mode == 0 ? (double)v32 : v64;
        }
      }
    }
  }
}

/* A4: Arellano backprojection, backprojection_cuda.cu:10-115.
 * transient [B][T][S][L], scan [B][3][S], laser [B][3][L], origins
 * [B][3][So] / [B][3][Lo] with So in {0,S}, Lo in {0,L}.  The reference's
 * interpolation weight is identically 0 because of an int truncation
 * (:101-103); the restatement keeps that. */
void oracle_arellano(const double* tr, int B, int T, int S, int L,
                     const float* scan, const float* laser,
                     const float* sorigin, int So, const float* lorigin,
                     int Lo, int D, int H, int W, const float* vs,
                     const float* ve, float w, float off, int mode,
                     double* out) {
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
  for (int b = 0; b < B; ++b) {
    for (int v = 0; v < H * W; ++v) {
      const int xi = v % W, yi = v / W;
      double* col64 = (double*)calloc((size_t)D, sizeof(double));
      float* col32 = (float*)calloc((size_t)D, sizeof(float));
      const float xf = inclusive32(xi, W, vs[0], ve[0]);   /* :29-30 */
      const float yf = inclusive32(yi, H, vs[1], ve[1]);   /* :31-32 */
      for (int s = 0; s < S; ++s) {
        const float sp[3] = {scan[(size_t)b * 3 * S + s],
                             scan[(size_t)b * 3 * S + S + s],
                             scan[(size_t)b * 3 * S + 2 * S + s]};
        float sd = 0.f;
        if (So > 0)
          sd = dist32(sp[0], sp[1], sp[2], sorigin[(size_t)b * 3 * So + s],
                      sorigin[(size_t)b * 3 * So + So + s],
                      sorigin[(size_t)b * 3 * So + 2 * So + s]);
        for (int l = 0; l < L; ++l) {
          const float lp[3] = {laser[(size_t)b * 3 * L + l],
                               laser[(size_t)b * 3 * L + L + l],
                               laser[(size_t)b * 3 * L + 2 * L + l]};
          float ld = 0.f;
          if (Lo > 0)
            ld = dist32(lp[0], lp[1], lp[2], lorigin[(size_t)b * 3 * Lo + l],
                        lorigin[(size_t)b * 3 * Lo + Lo + l],
                        lorigin[(size_t)b * 3 * Lo + 2 * Lo + l]);
          const float cx = 0.5f * (sp[0] + lp[0]);        /* :56-60 */
          const float cy = 0.5f * (sp[1] + lp[1]);
          const float cz = 0.5f * (sp[2] + lp[2]);
          float nx = sp[0] - lp[0], ny = sp[1] - lp[1], nz = sp[2] - lp[2];
          float ls = nx * nx;
          ls = ls + ny * ny;
          ls = ls + nz * nz;
          const float len = sqrtf(ls);                    /* :67 */
          if (len > 1.e-6f) {                             /* :68-70 */
            const float inv = 1.f / len;
            nx = inv * nx; ny = inv * ny; nz = inv * nz;
          }
          const float rsx = xf - cx, rsy = yf - cy, rsz = vs[2] - cz; /* :72 */
          const float rex = xf - cx, rey = yf - cy, rez = ve[2] - cz; /* :74 */
          for (int ti = 0; ti < T; ++ti) {
            const double val = tr[(((size_t)b * T + ti) * S + s) * L + l];
            const float value = (float)val;
            if (!(value > 0)) continue;                   /* :77 */
            float tc = (float)ti * w;                     /* :79-80 */
            tc = tc + off;
            tc = tc - sd;
            tc = tc - ld;
            if (!(tc > len)) continue;                    /* :81 */
            const float r0 = tc / 2.f;                    /* :83 */
            const float h = tc / 2.f;
            float r1s = h * h;
            r1s = r1s - (len * len) / 4.f;                /* :84 */
            const float r1 = sqrtf(r1s);
            const float sc = r1 / r0;                     /* :85 */
            float dps = rsx * nx;                         /* :87 dot */
            dps = dps + rsy * ny;
            dps = dps + rsz * nz;
            const float ks = (sc - 1.f) * dps;
            const float stx = rsx + ks * nx, sty = rsy + ks * ny, stz = rsz + ks * nz;
            float dpe = rex * nx;                         /* :88 */
            dpe = dpe + rey * ny;
            dpe = dpe + rez * nz;
            const float ke = (sc - 1.f) * dpe;
            const float etx = rex + ke * nx, ety = rey + ke * ny, etz = rez + ke * nz;
            float dx = etx - stx, dy = ety - sty, dz = etz - stz; /* :89 */
            float rl = dx * dx;
            rl = rl + dy * dy;
            rl = rl + dz * dz;
            const float rlen = sqrtf(rl);                 /* :90 */
            const float ir = 1.f / rlen;                  /* :91 */
            dx = ir * dx; dy = ir * dy; dz = ir * dz;
            float v0 = dx * stx;                          /* :93 */
            v0 = v0 + dy * sty;
            v0 = v0 + dz * stz;
            float ss = stx * stx;
            ss = ss + sty * sty;
            ss = ss + stz * stz;
            float v1 = v0 * v0 - (ss - r1 * r1);          /* :94 */
            if (!(v1 > 0)) continue;
            v1 = sqrtf(v1);
            float d = -v0 - v1;
            if (d < 0) d = -v0 + v1;                      /* :97-100 */
            float bf = d / rlen;
            bf = bf * (float)D;
            bf = bf + 0.5f;
            const int binf = (int)bf;                     /* :101 truncation */
            const int bini = binf;                        /* :102 */
            if (bini >= 0 && bini < D) {                  /* :104-106 */
              col32[bini] += value;
              col64[bini] += val;
            }
            /* :107-109 adds alpha*value with alpha == 0 */
          }
        }
      }
      for (int z = 0; z < D; ++z)
        out[(((size_t)b * D + z) * H + yi) * W + xi] = mode == 0 ? (double)col32[z] : col64[z];
      free(col64);
      free(col32);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* S1: Gaussian RBF splat, rbf2volume_cuda.cu:15-29,49-93.                   */
/* rbf [B][N][4+A], out [B][1+A][D][H][W].                                   */
/* trunc > 0: only RBFs with squared distance <= (trunc*sigma)^2 add.        */
/* mode 0: FP32 (expf), 1: FP64 exp on FP32 exponent inputs, 2: pure FP64.   */
/* ------------------------------------------------------------------------ */
static inline float splat_coord32(int i, int n, float a, float b) {
  /* rbf2volume_cuda.cu:58-60 */
  if (n == 1) return 0.5f * (a + b);
  float t = (b - a) * (float)i;
  t = t / (float)(n - 1);
  return a + t;
}

static inline float gauss32(float x, float y, float z, float xr, float yr,
                            float zr, float sg, float* r2_out) {
  /* rbf2volume_cuda.cu:20-28 */
  float r2 = (xr - x) * (xr - x);
  r2 = r2 + (yr - y) * (yr - y);
  r2 = r2 + (zr - z) * (zr - z);
  if (r2_out) *r2_out = r2;
  float e = -0.5f * r2;
  e = e / (sg * sg);
  return expf(e);
}

/* S1 over the whole grid, or (vox != NULL) only at the nvox listed flat
 * indices b * D*H*W + (z * H + y) * W + x (other outputs stay untouched). */
void oracle_splat_forward(const float* rbf, int B, int N, int A, int D,
                          int H, int W, const float* vs, const float* ve,
                          float trunc, int mode, double* out,
                          const long long* vox, long long nvox) {
  const int C = 4 + A;
  const size_t V = (size_t)D * H * W;
  const long long total = vox ? nvox : (long long)B * D * H;
#pragma omp parallel for schedule(dynamic, 1)
  for (long long job = 0; job < total; ++job) {
    const long long flat = vox ? vox[job] : -1;
    const int b = vox ? (int)(flat / (long long)V) : (int)(job / ((long long)D * H));
    const int zi = vox ? (int)((flat % (long long)V) / ((long long)H * W))
                       : (int)((job / H) % D);
    const int y_lo = vox ? (int)((flat % ((long long)H * W)) / W) : (int)(job % H);
    const int x_lo = vox ? (int)(flat % W) : 0;
    const int x_hi = vox ? x_lo + 1 : W;
    {
      double* a64 = (double*)malloc(sizeof(double) * (A > 0 ? A : 1));
      float* a32 = (float*)malloc(sizeof(float) * (A > 0 ? A : 1));
      for (int yi = y_lo; yi <= y_lo; ++yi) {
        for (int xi = x_lo; xi < x_hi; ++xi) {
          const float x = splat_coord32(xi, W, vs[0], ve[0]);
          const float y = splat_coord32(yi, H, vs[1], ve[1]);
          const float z = splat_coord32(zi, D, vs[2], ve[2]);
          float s32 = 0.f;
          double s64 = 0.0;
          for (int a = 0; a < A; ++a) { a32[a] = 0.f; a64[a] = 0.0; }
          for (int i = 0; i < N; ++i) {
            const float* r = rbf + ((size_t)b * N + i) * C;
            float r2;
            float g = gauss32(x, y, z, r[0], r[1], r[2], r[3], &r2);
            if (trunc > 0.f) {
              const float lim = trunc * r[3];
              if (r2 > lim * lim) continue;
            }
            double g64;
            if (mode == 0) {
              g64 = g;
            } else if (mode == 1) {
              float e = -0.5f * r2;
              e = e / (r[3] * r[3]);
              g64 = exp((double)e);
            } else {
              double dx = (double)r[0] - x, dy = (double)r[1] - y, dz = (double)r[2] - z;
              g64 = exp(-0.5 * (dx * dx + dy * dy + dz * dz) / ((double)r[3] * r[3]));
            }
            for (int a = 0; a < A; ++a) {
              a32[a] += g * r[4 + a];
              a64[a] += g64 * (double)r[4 + a];
            }
            s32 += g;
            s64 += g64;
          }
          const size_t vi = ((size_t)zi * H + yi) * W + xi;
          double* ob = out + (size_t)b * (1 + A) * V;
          ob[vi] = mode == 0 ? (double)s32 : s64;
          if (mode == 0) {
            float den = s32 < 1.e-8f ? 1.f : s32;
            for (int a = 0; a < A; ++a) ob[(size_t)(1 + a) * V + vi] = (double)(a32[a] / den);
          } else {
            double den = s64 < 1.e-8 ? 1.0 : s64;
            for (int a = 0; a < A; ++a) ob[(size_t)(1 + a) * V + vi] = a64[a] / den;
          }
        }
      }
      free(a64);
      free(a32);
    }
  }
}

/* S2: splat gradient, rbf2volume_cuda.cu:31-47,123-201.
 * vol and vgrad [B][1+A][D][H][W] (vol is the forward output), out
 * [B][N][4+A].  Sums over voxels within the truncation radius (all voxels
 * when trunc <= 0) in z, y, x order, in FP32 (mode 0) or FP64 (mode 1/2;
 * mode 2 recomputes the Gaussian in FP64). */
void oracle_splat_grad(const float* rbf, int B, int N, int A,
                       const double* vol, const double* vgrad, int D, int H,
                       int W, const float* vs, const float* ve, float trunc,
                       int mode, double* out) {
  const int C = 4 + A;
  const size_t V = (size_t)D * H * W;
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
  for (int b = 0; b < B; ++b) {
    for (int i = 0; i < N; ++i) {
      const float* r = rbf + ((size_t)b * N + i) * C;
      const float xr = r[0], yr = r[1], zr = r[2], sg = r[3];
      double acc64[4 + 8 + 64];
      float acc32[4 + 8 + 64];
      for (int c = 0; c < C; ++c) { acc64[c] = 0; acc32[c] = 0; }
      const double* vb = vol + (size_t)b * (1 + A) * V;
      const double* gb = vgrad + (size_t)b * (1 + A) * V;
      /* With truncation only voxels inside the support box can pass the
       * radius test; the box is widened by two grid steps so it contains
       * every voxel that passes, and the loop order is unchanged. */
      int lo[3] = {0, 0, 0}, hi[3] = {W - 1, H - 1, D - 1};
      if (trunc > 0.f) {
        const int n3[3] = {W, H, D};
        const float c3[3] = {xr, yr, zr};
        for (int d = 0; d < 3; ++d) {
          if (n3[d] < 2) continue;
          const double step = ((double)ve[d] - vs[d]) / (n3[d] - 1);
          if (step == 0.0) continue;
          const double rad = (double)trunc * sg;
          double a = (c3[d] - rad - vs[d]) / step, e = (c3[d] + rad - vs[d]) / step;
          if (a > e) { const double t = a; a = e; e = t; }
          a = floor(a) - 2;
          e = ceil(e) + 2;
          lo[d] = a < 0 ? 0 : (a > n3[d] - 1 ? n3[d] : (int)a);
          hi[d] = e > n3[d] - 1 ? n3[d] - 1 : (e < 0 ? -1 : (int)e);
        }
      }
      for (int zi = lo[2]; zi <= hi[2]; ++zi)
        for (int yi = lo[1]; yi <= hi[1]; ++yi)
          for (int xi = lo[0]; xi <= hi[0]; ++xi) {
            const size_t vi = ((size_t)zi * H + yi) * W + xi;
            const float x = splat_coord32(xi, W, vs[0], ve[0]);
            const float y = splat_coord32(yi, H, vs[1], ve[1]);
            const float z = splat_coord32(zi, D, vs[2], ve[2]);
            float r2;
            const float wgt = gauss32(x, y, z, xr, yr, zr, sg, &r2);
            if (trunc > 0.f) {
              const float lim = trunc * sg;
              if (r2 > lim * lim) continue;
            }
            if (mode == 0) {
              const float si = 1.f / sg;                          /* :38 */
              const float dxr = (((x - xr) * si) * si) * wgt;     /* :39 */
              const float dyr = (((y - yr) * si) * si) * wgt;
              const float dzr = (((z - zr) * si) * si) * wgt;
              const float dsr = (((r2 * si) * si) * si) * wgt;    /* :42-46 */
              const float dl0 = (float)gb[vi];
              if (A == 0) {                                       /* :159-163 */
                acc32[0] += dl0 * dxr;
                acc32[1] += dl0 * dyr;
                acc32[2] += dl0 * dzr;
                acc32[3] += dl0 * dsr;
              } else {
                float v0 = (float)vb[vi];
                if (v0 < 1.e-8f) v0 = 1.f;                        /* :165-169 */
                const float inv = 1.f / v0;
                float sum = 0.f;
                for (int a = 0; a < A; ++a) {
                  const float at = (float)vb[(size_t)(1 + a) * V + vi];
                  const float dla = (float)gb[(size_t)(1 + a) * V + vi];
                  sum += dla * (r[4 + a] - at);                   /* :176 */
                  acc32[4 + a] += (dla * wgt) * inv;              /* :178 */
                }
                sum = sum * inv;                                  /* :180 */
                acc32[0] += dxr * (dl0 + sum);                    /* :181-184 */
                acc32[1] += dyr * (dl0 + sum);
                acc32[2] += dzr * (dl0 + sum);
                acc32[3] += dsr * (dl0 + sum);
              }
            } else {
              double g;
              double dx = (double)x - xr, dy = (double)y - yr, dz = (double)z - zr;
              double rr = dx * dx + dy * dy + dz * dz;
              if (mode == 2)
                g = exp(-0.5 * rr / ((double)sg * sg));
              else
                g = exp((double)(-0.5f * r2 / (sg * sg)));
              const double si = 1.0 / sg;
              const double dxr = dx * si * si * g, dyr = dy * si * si * g,
                           dzr = dz * si * si * g, dsr = rr * si * si * si * g;
              const double dl0 = gb[vi];
              if (A == 0) {
                acc64[0] += dl0 * dxr;
                acc64[1] += dl0 * dyr;
                acc64[2] += dl0 * dzr;
                acc64[3] += dl0 * dsr;
              } else {
                double v0 = vb[vi];
                if (v0 < 1.e-8) v0 = 1.0;
                const double inv = 1.0 / v0;
                double sum = 0.0;
                for (int a = 0; a < A; ++a) {
                  const double at = vb[(size_t)(1 + a) * V + vi];
                  const double dla = gb[(size_t)(1 + a) * V + vi];
                  sum += dla * ((double)r[4 + a] - at);
                  acc64[4 + a] += dla * g * inv;
                }
                sum *= inv;
                acc64[0] += dxr * (dl0 + sum);
                acc64[1] += dyr * (dl0 + sum);
                acc64[2] += dzr * (dl0 + sum);
                acc64[3] += dsr * (dl0 + sum);
              }
            }
          }
      for (int c = 0; c < C; ++c)
        out[((size_t)b * N + i) * C + c] = mode == 0 ? (double)acc32[c] : acc64[c];
    }
  }
}

/* ------------------------------------------------------------------------ */
/* R1 decisions under nvcc's default contraction (--fmad=true), against the */
/* uncontracted reading of volume_cuda.cu.  nvcc fuses a product into the   */
/* following sum where the source writes one:                                */
/*   :23  tc = ti * bin_width + bin_offset   -> fmaf(ti, w, off)             */
/*   :46  z_hat = (z - o) * s - 0.5          -> fmaf(z - o, s, -0.5)         */
/* (depth^2 is built from powf calls, and x from a product then a quotient, */
/* so neither has a product feeding a sum in the source.)                    */
/* Counts, over rows [i0, i1) of the scan, every bin and every column:       */
/*   out[0] samples active in the uncontracted reading                        */
/*   out[1] samples whose activity differs between the readings               */
/*   out[2] samples active in both whose floor index differs                  */
/* ------------------------------------------------------------------------ */
static inline int r1_decide(float tc, float wx, float wy, float wz, float x, float y,
                            float oz, float s, int Z, int fused, int* zi_out) {
  float t2 = sq32(tc);
  float dx = wx - x, dy = wy - y;
  float d2 = t2 - sq32(dx);
  d2 = d2 - sq32(dy);
  if (!(d2 > 0.f)) return 0;
  float depth = sqrtf(d2);
  float z = wz + depth;
  float zh;
  if (fused) {
    zh = fmaf(z - oz, s, -0.5f);
  } else {
    zh = (z - oz) * s;
    zh = zh - 0.5f;
  }
  const float fl = floorf(zh);
  if (!(fl >= 0.f && fl + 1.f < (float)Z)) return 0;
  *zi_out = (int)fl;
  return 1;
}

void oracle_r1_contraction_flips(const float* mp, int N, int T, int X, int Y, int Z,
                                 const float* origin, const float* size, float w,
                                 float off, int i0, int i1, long long* out) {
  const float s = (float)Z / size[2];
  long long act = 0, aflip = 0, zflip = 0;
#pragma omp parallel for collapse(2) schedule(dynamic, 4) reduction(+ : act, aflip, zflip)
  for (int i = i0; i < i1; ++i) {
    for (int ti = 0; ti < T; ++ti) {
      const float wx = mp[i], wy = mp[N + i], wz = mp[2 * N + i];
      const float tc = r1_tc32(ti, w, off);
      const float tcf = fmaf((float)ti, w, off);
      for (int xi = 0; xi < X; ++xi) {
        const float x = r1_x32(xi, origin[0], size[0], X);
        for (int yi = 0; yi < Y; ++yi) {
          const float y = r1_x32(yi, origin[1], size[1], Y);
          int za = 0, zb = 0;
          const int a = r1_decide(tc, wx, wy, wz, x, y, origin[2], s, Z, 0, &za);
          const int b = r1_decide(tcf, wx, wy, wz, x, y, origin[2], s, Z, 1, &zb);
          act += a;
          aflip += (a != b);
          zflip += (a && b && za != zb);
        }
      }
    }
  }
  out[0] = act;
  out[1] = aflip;
  out[2] = zflip;
}

/* ------------------------------------------------------------------------ */
/* The reference's double dispatch (AT_DISPATCH_FLOATING_TYPES with          */
/* scalar_t = double: volume_cuda.cu:85,175; backprojection_cuda.cu:215).    */
/* Geometry arguments stay float, and some steps keep float arithmetic:      */
/*   tc = ti * bin_width + bin_offset is a float expression, widened         */
/*        (volume_cuda.cu:23,137)                                             */
/*   floorf(z_hat) narrows z_hat to float before the floor (:47,:147)        */
/*   ceilf(tf) narrows tf to float (:132-134)                                 */
/*   A2 takes its points as float3 and measures float distances              */
/*        (backprojection_cuda.cu:167-183); only the bin arithmetic and the   */
/*        sum are double (:185-190), and floorf narrows again.                */
/* pow(x, 2) and pow(x, 4) are restated as products (CUDA's pow: unpinned).   */
/* ------------------------------------------------------------------------ */
static inline double r1_x64d(int xi, float o, float sz, int n) {
  /* volume_cuda.cu:35  scalar_t(o) + (xi + scalar_t(0.5)) * scalar_t(sz) / scalar_t(n) */
  double a = ((double)xi + 0.5) * (double)sz;
  a = a / (double)n;
  return (double)o + a;
}

typedef struct {
  int active;
  int zi;
  double alpha;
  double f2;
} r1_sample_d;

static inline r1_sample_d r1_sample_f64(double tc, double wx, double wy, double wz,
                                        double x, double y, double o, double s, int Z) {
  r1_sample_d r = {0, 0, 0.0, 0.0};
  const double t2 = tc * tc;                /* :40 left to right */
  const double dx = wx - x, dy = wy - y;
  double d2 = t2 - dx * dx;
  d2 = d2 - dy * dy;
  if (!(d2 > 0.0)) return r;                /* :41 */
  const double depth = sqrt(d2);
  const double z = wz + depth;              /* :45 */
  double zh = (z - o) * s;
  zh = zh - 0.5;                            /* :47 */
  const int zi = (int)floorf((float)zh);    /* :48 floorf narrows to float */
  if (!(zi >= 0 && zi + 1 < Z)) return r;   /* :50 */
  r.active = 1;
  r.zi = zi;
  r.alpha = zh - (double)zi;                /* :52 */
  const double q = -depth / t2;
  const double q2 = q * q;
  r.f2 = q2 * q2;                           /* :58 */
  return r;
}

/* R1 in double, transient rows [i0, i1) of [N][T]. */
void oracle_r1_forward_f64(const double* vol, int X, int Y, int Z, const double* mp, int N,
                           int T, const float* origin, const float* size, float w, float off,
                           double* out, int i0, int i1) {
  const double o = (double)origin[2];
  const double s = (double)Z / (double)size[2];   /* :29 */
#pragma omp parallel for collapse(2) schedule(dynamic, 16)
  for (int i = i0; i < i1; ++i) {
    for (int ti = 0; ti < T; ++ti) {
      const double wx = mp[i], wy = mp[N + i], wz = mp[2 * N + i];
      const double tc = (double)r1_tc32(ti, w, off);   /* :23 float, widened */
      double value = 0.0;
      for (int xi = 0; xi < X; ++xi) {
        const double x = r1_x64d(xi, origin[0], size[0], X);
        for (int yi = 0; yi < Y; ++yi) {
          const double y = r1_x64d(yi, origin[1], size[1], Y);
          const r1_sample_d sm = r1_sample_f64(tc, wx, wy, wz, x, y, o, s, Z);
          if (!sm.active) continue;
          const double* col = vol + ((size_t)xi * Y + yi) * Z;
          const double f1 = (1.0 - sm.alpha) * col[sm.zi] + sm.alpha * col[sm.zi + 1]; /* :55 */
          value = value + f1 * sm.f2;                                                  /* :60 */
        }
      }
      out[(size_t)(i - i0) * T + ti] = value;
    }
  }
}

/* A1 in double as the exact transpose of oracle_r1_forward_f64: each sample
 * adds (1 - alpha) f2 G into zi and alpha f2 G into zi + 1, scan points and
 * bins in order.  Columns listed in `cols` (pairs xi, yi), all when NULL. */
void oracle_a1_transpose_f64(const double* G, const double* mp, int N, int T, int X, int Y,
                             int Z, const float* origin, const float* size, float w, float off,
                             const int* cols, int ncols, double* out) {
  const double o = (double)origin[2];
  const double s = (double)Z / (double)size[2];
  const int nc = cols ? ncols : X * Y;
#pragma omp parallel for schedule(dynamic, 1)
  for (int k = 0; k < nc; ++k) {
    const int xi = cols ? cols[2 * k] : k / Y;
    const int yi = cols ? cols[2 * k + 1] : k % Y;
    double* col = out + ((size_t)xi * Y + yi) * Z;
    for (int z = 0; z < Z; ++z) col[z] = 0.0;
    const double x = r1_x64d(xi, origin[0], size[0], X);
    const double y = r1_x64d(yi, origin[1], size[1], Y);
    for (int i = 0; i < N; ++i) {
      const double wx = mp[i], wy = mp[N + i], wz = mp[2 * N + i];
      for (int ti = 0; ti < T; ++ti) {
        const double tc = (double)r1_tc32(ti, w, off);
        const r1_sample_d sm = r1_sample_f64(tc, wx, wy, wz, x, y, o, s, Z);
        if (!sm.active) continue;
        const double wg = sm.f2 * G[(size_t)i * T + ti];
        col[sm.zi] += (1.0 - sm.alpha) * wg;
        col[sm.zi + 1] += sm.alpha * wg;
      }
    }
  }
}

/* The reference's window-form A1 in double (volume_cuda.cu:95-160). */
void oracle_a1_reference_f64(const double* G, const double* mp, int N, int T, int X, int Y,
                             int Z, const float* origin, const float* size, float w, float off,
                             const int* cols, int ncols, double* out) {
  const double o = (double)origin[2];
  const double s = (double)Z / (double)size[2];   /* :112 */
  const double dz = (double)size[2] / (double)Z;  /* :114 */
  const int nc = cols ? ncols : X * Y;
#pragma omp parallel for schedule(dynamic, 1)
  for (int k = 0; k < nc; ++k) {
    const int xi = cols ? cols[2 * k] : k / Y;
    const int yi = cols ? cols[2 * k + 1] : k % Y;
    const double x = r1_x64d(xi, origin[0], size[0], X);   /* :107 */
    const double y = r1_x64d(yi, origin[1], size[1], Y);   /* :108 */
    for (int zi = 0; zi < Z; ++zi) {
      const double z = r1_x64d(zi, origin[2], size[2], Z); /* :109 */
      double value = 0.0;
      for (int i = 0; i < N; ++i) {
        const double wx = mp[i], wy = mp[N + i], wz = mp[2 * N + i];
        const double zz = zi > 0 ? dz : 0.0;                         /* :123 */
        const double ax = x - wx, ay = y - wy;
        const double dps = ax * ax + ay * ay;                        /* :124 */
        const double a = (z - zz) - wz, b = z - wz, c = (z + dz) - wz;
        const double ds = sqrt(dps + a * a);                         /* :125 */
        const double dm = sqrt(dps + b * b);                         /* :126 */
        const double de = sqrt(dps + c * c);                         /* :127 */
        const double tfs = (ds - (double)off) / (double)w;           /* :129-131 */
        const double tfm = (dm - (double)off) / (double)w;
        const double tfe = (de - (double)off) / (double)w;
        int ts = (int)ceilf((float)tfs), tm = (int)ceilf((float)tfm),
            te = (int)ceilf((float)tfe);                             /* ceilf narrows */
        ts = ts < 0 ? 0 : (ts > T ? T : ts);                         /* :132-134 */
        tm = tm < 0 ? 0 : (tm > T ? T : tm);
        te = te < 0 ? 0 : (te > T ? T : te);
        for (int ti = ts; ti < te; ++ti) {                           /* :135 */
          const double tc = (double)r1_tc32(ti, w, off);             /* :137 */
          const double t2 = tc * tc;
          const double d2 = t2 - dps;                                /* :139 */
          if (!(d2 > 0.0)) continue;
          const double depth = sqrt(d2);
          const double zc = wz + depth;                              /* :143 */
          double zh = (zc - o) * s;
          zh = zh - 0.5;                                             /* :145 */
          const int zf = (int)floorf((float)zh);                     /* :146 floorf narrows */
          const double alpha = zh - (double)zf;                      /* :148 */
          const double df1 = (ti >= tm) ? (1.0 - alpha) : alpha;     /* :150 */
          const double q = -depth / t2;
          const double q2 = q * q;
          const double f2 = q2 * q2;                                 /* :152 */
          value = value + (df1 * f2) * G[(size_t)i * T + ti];        /* :154 */
        }
      }
      out[((size_t)xi * Y + yi) * Z + zi] = value;
    }
  }
}

/* A2 in double (backprojection_cuda.cu:152-194 with scalar_t = double):
 * float points and distances, double bin arithmetic and sum. */
void oracle_velten_confocal_f64(const double* tr, int B, int T, int S, const double* scan,
                                const double* sorigin, int So, int D, int H, int W,
                                const float* vs, const float* ve, float w, float off,
                                double* out) {
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
  for (int b = 0; b < B; ++b) {
    for (int z = 0; z < D; ++z) {
      for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
          const float qx = inclusive32(x, W, vs[0], ve[0]);  /* :167-170 */
          const float qy = inclusive32(y, H, vs[1], ve[1]);
          const float qz = inclusive32(z, D, vs[2], ve[2]);
          double v = 0.0;
          for (int s = 0; s < S; ++s) {
            const double* p = scan + ((size_t)b * S + s) * 3;
            const float px = (float)p[0], py = (float)p[1], pz = (float)p[2]; /* :173-176 */
            const double dsv = (double)dist32(px, py, pz, qx, qy, qz);       /* :177 */
            double dso = 0.0;
            if (So > 0) {
              const double* q = sorigin + ((size_t)b * So + s) * 3;
              dso = (double)dist32(px, py, pz, (float)q[0], (float)q[1], (float)q[2]); /* :178-183 */
            }
            const double dist = 2.0 * dso + 2.0 * dsv;                        /* :185 */
            const double tf = (dist - (double)off) / (double)w + 0.5;
            const int ti = (int)floorf((float)tf);                            /* :187 */
            if (ti >= 0 && ti < T) v += tr[((size_t)b * T + ti) * S + s];    /* :188-190 */
          }
          out[(((size_t)b * D + z) * H + y) * W + x] = v;
        }
      }
    }
  }
}
