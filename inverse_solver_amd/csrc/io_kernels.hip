// 2D input/output projection kernels (see io_kernels.hpp).
#include "io_kernels.hpp"

namespace cbs {

__device__ __forceinline__ int wrap(int i, int n) {
  i %= n;
  return i < 0 ? i + n : i;
}

// buf (conv grid, ifftshift order) = ifftshift(padd_field2conv(src)).
// ifftshift: out[d] = in[(d + n/2) mod n].
__global__ void io_pad_in(const float2* __restrict__ src, float2* __restrict__ buf, IoGeom g, int nb) {
  const long long n2 = (long long)g.cx * g.cy;
  const long long tot_ = n2 * nb;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < tot_;
       i += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(i / n2);
    const long long j = i - (long long)r * n2;
    const int X = (int)(j % g.cx), Y = (int)(j / g.cx);
    const int xp = (X + g.cx / 2) % g.cx - g.pad_x;
    const int yp = (Y + g.cy / 2) % g.cy - g.pad_y;
    float2 v = make_float2(0.f, 0.f);
    if (xp >= 0 && xp < g.fx && yp >= 0 && yp < g.fy) v = src[(long long)r * g.fx * g.fy + (long long)yp * g.fx + xp];
    buf[i] = v;
  }
}

// buf3[r][z][p] = spec2[r][p] * refocus_centred[z][kc(p)] * scale, for z < rz.
__global__ void io_refocus(const float2* __restrict__ spec2, const float2* __restrict__ refocus,
                           float2* __restrict__ buf3, const int* __restrict__ kcx, const int* __restrict__ kcy,
                           IoGeom g, int nb, float scale) {
  const long long n2 = (long long)g.cx * g.cy;
  const long long tot = n2 * g.rz * nb;
  const long long tot_ = tot;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < tot_;
       i += (long long)gridDim.x * blockDim.x) {
    const long long j = i % n2;
    const long long rz_ = i / n2;
    const int z = (int)(rz_ % g.rz);
    const int r = (int)(rz_ / g.rz);
    const int px = (int)(j % g.cx), py = (int)(j / g.cx);
    const float2 k = refocus[((long long)z * g.cy + kcy[py]) * g.cx + kcx[px]];
    buf3[i] = cscale(cmul(spec2[(long long)r * n2 + j], k), scale);
  }
}

// u_in[r][z][yr][xr] = fftshift(buf3)[ri_y + yr][ri_x + xr]; fftshift: out[d] = in[(d - n/2) mod n].
__global__ void io_crop_in(const float2* __restrict__ buf3, float2* __restrict__ u_in, IoGeom g, int nb) {
  const long long nri = (long long)g.rx * g.ry * g.rz;
  const long long tot_ = nri * nb;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < tot_;
       i += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(i / nri);
    const long long j = i - (long long)r * nri;
    const int xr = (int)(j % g.rx);
    const long long yz = j / g.rx;
    const int yr = (int)(yz % g.ry), z = (int)(yz / g.ry);
    const int X = wrap(g.ri_x + xr - g.cx / 2, g.cx);
    const int Y = wrap(g.ri_y + yr - g.cy / 2, g.cy);
    u_in[i] = buf3[(((long long)r * g.rz + z) * g.cy + Y) * g.cx + X];
  }
}

// buf3[r][z] (ifftshift order) = ifftshift(padd_RI2conv(u * Vt * dV)), u = Field[ROI] + u_in,
// Vt = V + i eps on the RI grid.
__global__ void io_emitter(const float2* __restrict__ field, const float2* __restrict__ u_in,
                           const float2* __restrict__ V, float2* __restrict__ buf3, IoGeom g, Consts c, int ox, int oy,
                           int oz, long long rhs_stride, int nb, float dV) {
  const long long n2 = (long long)g.cx * g.cy;
  const long long tot = n2 * g.rz * nb;
  const long long tot_ = tot;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < tot_;
       i += (long long)gridDim.x * blockDim.x) {
    const long long j = i % n2;
    const long long rz_ = i / n2;
    const int z = (int)(rz_ % g.rz);
    const int r = (int)(rz_ / g.rz);
    const int X = (int)(j % g.cx), Y = (int)(j / g.cx);
    const int xr = (X + g.cx / 2) % g.cx - g.ri_x;
    const int yr = (Y + g.cy / 2) % g.cy - g.ri_y;
    float2 e = make_float2(0.f, 0.f);
    if (xr >= 0 && xr < g.rx && yr >= 0 && yr < g.ry) {
      const long long gs = ((long long)(z + oz) * c.ny + (yr + oy)) * c.nx + (xr + ox);
      const long long gf = ((long long)(z + oz) * c.ny + state_row(c, yr + oy)) * c.nx + (xr + ox);
      const long long gr = (((long long)r * g.rz + z) * g.ry + yr) * g.rx + xr;
      const float2 u = cadd(field[(long long)r * rhs_stride + gf], u_in[gr]);
      const float2 v = V[gs];
      const float2 vt = make_float2(v.x * dV, (v.y + c.eps) * dV);
      e = cmul(u, vt);
  }
  buf3[i] = e;
  }
}

// out[r][p] = sum_z buf3[r][z][p] * K_centred[z][kc(p)], FP64 accumulation.
__global__ void io_zsum(const float2* __restrict__ buf3, const float2* __restrict__ ktrans,
                        const float2* __restrict__ kref, float2* __restrict__ out_t, float2* __restrict__ out_r,
                        const int* __restrict__ kcx, const int* __restrict__ kcy, IoGeom g, int nb) {
  const long long n2 = (long long)g.cx * g.cy;
  const long long tot_ = n2 * nb;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < tot_;
       i += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(i / n2);
    const long long j = i - (long long)r * n2;
    const int px = (int)(j % g.cx), py = (int)(j / g.cx);
    const long long kj = (long long)kcy[py] * g.cx + kcx[px];
    double tr = 0.0, ti = 0.0, rr = 0.0, ri = 0.0;
    // z in chunks of ZC planes whose loads are all issued before the first is used (one
    // output per thread leaves few waves per CU for a single rhs); the sums keep z order
    constexpr int ZC = 8;
    const float2* eb = buf3 + (long long)r * g.rz * n2 + j;
    for (int z0 = 0; z0 < g.rz; z0 += ZC) {
      float2 e[ZC], kt[ZC], kr[ZC];
#pragma unroll
      for (int q = 0; q < ZC; ++q) {
        const int z = z0 + q;
        if (z < g.rz) {
          e[q] = eb[(long long)z * n2];
          if (out_t) kt[q] = ktrans[(long long)z * n2 + kj];
          if (out_r) kr[q] = kref[(long long)z * n2 + kj];
        }
      }
#pragma unroll
      for (int q = 0; q < ZC; ++q) {
        if (z0 + q >= g.rz) break;
        if (out_t) {
          tr += (double)e[q].x * kt[q].x - (double)e[q].y * kt[q].y;
          ti += (double)e[q].x * kt[q].y + (double)e[q].y * kt[q].x;
        }
        if (out_r) {
          rr += (double)e[q].x * kr[q].x - (double)e[q].y * kr[q].y;
          ri += (double)e[q].x * kr[q].y + (double)e[q].y * kr[q].x;
        }
      }
    }
  if (out_t) out_t[i] = make_float2((float)tr, (float)ti);
  if (out_r) out_r[i] = make_float2((float)rr, (float)ri);
  }
}

// out[r][yr][xr] = scale * fftshift(buf2)[pad_y + yr][pad_x + xr]  (crop_conv2field).
__global__ void io_crop_out(const float2* __restrict__ buf2, float2* __restrict__ out, IoGeom g, int nb, float scale) {
  const long long nf = (long long)g.fx * g.fy;
  const long long tot_ = nf * nb;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < tot_;
       i += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(i / nf);
    const long long j = i - (long long)r * nf;
    const int xr = (int)(j % g.fx), yr = (int)(j / g.fx);
    const int X = wrap(g.pad_x + xr - g.cx / 2, g.cx);
    const int Y = wrap(g.pad_y + yr - g.cy / 2, g.cy);
    out[i] = cscale(buf2[((long long)r * g.cy + Y) * g.cx + X], scale);
  }
}

// Centred pupil filter of fields on the field grid: the centring shifts cancel,
// fftshift(ifft2(ifftshift(P) fft2(ifftshift(x)))) = ifft2(ifftshift(P) fft2(x)) for any
// length n (the shift by n/2 multiplies the spectrum by exp(2 pi i k (n/2) / n) on the way
// in and by its conjugate on the way out), so the filter is field_sum (x = a [+ b]), the
// field-grid FFTs in place, and field_pupil_mul with the table of ifftshift(P) at each
// spectrum position (the 1/N of the inverse transform folded in).
__global__ void field_sum(const float2* a, const float2* __restrict__ b, float2* out,   // out may be a
                          long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    out[i] = b ? cadd(a[i], b[i]) : a[i];
}

__global__ void field_pupil_mul(float2* __restrict__ buf, const float* __restrict__ pupil, int nx, int ny, long long nf) {
  const long long n2 = (long long)nx * ny;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n2 * nf;
       i += (long long)gridDim.x * blockDim.x)
    buf[i] = cscale(buf[i], pupil[i % n2]);
}

}  // namespace cbs
