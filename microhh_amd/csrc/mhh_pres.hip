// mhh_pres.hip -- 2nd-order pressure projection (Pres_2, src/pres_2.cxx:64-422).
//
// Spectral layout.  The reference transforms with FFTW R2HC per k-slice along
// x then y (src/fft.cxx:338-394) and solves one tridiagonal system per
// halfcomplex entry.  Here the same eigen-decomposition is done with complex
// arithmetic: a real-to-complex transform in x (two real rows packed in one
// complex FFT), a complex transform in y, a Thomas solve per (kx, ky) mode
// with the reference's coefficients and boundary rows, and the inverse path.
// The solution p is the same up to rounding; only the intermediate
// representation differs.
//
// Memory layout of the spectrum (complex, in tmp1):
//     C[k][kx][j],  k < ktot, kx <= itot/2, j < jtot   (j contiguous)
// so that the y transforms read contiguous rows and the Thomas solve reads
// consecutive modes with consecutive threads.  The x transforms transpose
// through LDS to write/read 2*RB consecutive j values per kx.
//
// FFT: Stockham autosort, radix 4/2/3/5 passes ping-ponging between two LDS
// buffers, one 64-wide wavefront per row, twiddles from a host-computed table.
#include <algorithm>
#include <cmath>
#include <vector>

#include "mhh_internal.h"
#include "mhh_fft.h"

namespace
{
    using namespace mhh;

    // Pres_2::input (src/pres_2.cxx:155-196): divergence of u/dt + ut into the
    // dense [k][j][i] array P.  The periodic neighbours of ut and vt at the east
    // and north edges are read by index wrap instead of through their halos.
    template<typename TF>
    __global__ void __launch_bounds__(BX*BY) pres_in_g(
            TF* __restrict__ P, const TF* __restrict__ u, const TF* __restrict__ v, const TF* __restrict__ w,
            const TF* __restrict__ ut, const TF* __restrict__ vt, const TF* __restrict__ wt,
            const TF* __restrict__ dzi, const TF* __restrict__ rho, const TF* __restrict__ rhoh,
            TF dxi, TF dyi, TF dti, Gd g)
    {
        const int i = blockIdx.x*BX + threadIdx.x;
        const int j = blockIdx.y*BY + threadIdx.y;
        const int k = blockIdx.z;
        if (i >= g.itot || j >= g.jtot) return;
        const int jj = g.icells, kk = g.ijcells;
        const int ijk = i+g.igc + (j+g.jgc)*jj + (k+g.kgc)*kk;
        const int ie = (i+1 == g.itot) ? ijk + 1 - g.itot : ijk + 1;          // east neighbour of ut
        const int jn = (j+1 == g.jtot) ? ijk + jj - g.jtot*jj : ijk + jj;     // north neighbour of vt
        const int kc = k + g.kgc;
        P[i + j*g.itot + k*g.itot*g.jtot] =
                  rho[kc] * ( (ut[ie] + u[ijk+1] * dti) - (ut[ijk] + u[ijk] * dti) ) * dxi
                + rho[kc] * ( (vt[jn] + v[ijk+jj] * dti) - (vt[ijk] + v[ijk] * dti) ) * dyi
                + ( rhoh[kc+1] * (wt[ijk+kk] + w[ijk+kk] * dti)
                  - rhoh[kc  ] * (wt[ijk   ] + w[ijk   ] * dti) ) * dzi[kc];
    }

    // Packed staging of the two half-spectra of a pair of real rows in n complex
    // slots: the DC (and, for even n, Nyquist) entries are real and share one
    // slot each as (row a, row b); every other kx holds (Xa, Xb) in two slots.
    __device__ __forceinline__ int stage_base(int kx, int n) { return (n & 1) ? 2*kx - 1 : 2*kx; }

    constexpr int MAXQ = 9;   // (n/2 + 1)/64 rounded up for the largest fast plan (n = 1024)

    template<typename TF>
    __device__ __forceinline__ void unpack_pair(cplx<TF> z, cplx<TF> c, cplx<TF>& xa, cplx<TF>& xb)
    {
        // spectra of the real rows a, b from Z = FFT(a + i b): Xa = (Z[k] + conj(Z[-k]))/2,
        // Xb = (Z[k] - conj(Z[-k]))/(2i)
        xa.x = (z.x + c.x)*TF(0.5); xa.y = (z.y - c.y)*TF(0.5);
        xb.x = (z.y + c.y)*TF(0.5); xb.y = (c.x - z.x)*TF(0.5);
    }

    template<typename TF>
    __device__ __forceinline__ void stage_store(cplx<TF>* S, int kx, int n, cplx<TF> xa, cplx<TF> xb)
    {
        if (kx == 0) { cplx<TF> d; d.x = xa.x; d.y = xb.x; S[0] = d; }
        else if (2*kx == n) { cplx<TF> d; d.x = xa.x; d.y = xb.x; S[1] = d; }
        else { const int b = stage_base(kx, n); S[b] = xa; S[b+1] = xb; }
    }

    // Forward x transform: RB complex rows per block, each packing two real
    // rows of P at level k; writes C[k][kx][j] for kx <= itot/2.
    template<typename TF>
    __global__ void __launch_bounds__(256) fft_x_fwd_g(const TF* __restrict__ P, cplx<TF>* __restrict__ C,
                                                       Plan plan, const double2* __restrict__ tw, int RB, Gd g)
    {
        using T = cplx<TF>;
        extern __shared__ unsigned char smem_raw[];
        T* tws = reinterpret_cast<T*>(smem_raw);
        const int n = plan.n, nx = n/2 + 1, M = g.jtot, re = row_elems(plan);
        T* rows = tws + n;
        const int row = threadIdx.y, lane = threadIdx.x;
        const int k = blockIdx.y;
        const int j0 = blockIdx.x*2*RB;              // first real row of the block
        const int ja = j0 + 2*row, jb = ja + 1;
        T* A = rows + row*re;
        if (plan.fast) load_twiddles<TF>(tws, tw, n);
        const TF* Pk = P + static_cast<size_t>(k)*n*M;
        for (int x = lane; x < n; x += 64)
        {
            T z;
            z.x = (ja < M) ? Pk[x + ja*n] : TF(0);
            z.y = (jb < M) ? Pk[x + jb*n] : TF(0);
            A[x] = z;
        }
        __syncthreads();
        T* Z = fft_row<TF>(A, tws, tw, plan, -1, lane);
        // Stage the two half-spectra packed in n slots: in the row buffer itself
        // for the in-place engine (through registers), else in the free buffer.
        int soff;
        if (plan.fast)
        {
            T xs[MAXQ][2];
            #pragma unroll
            for (int q = 0; q < MAXQ; ++q)
            {
                const int kx = lane + 64*q;
                if (kx < nx) unpack_pair<TF>(Z[kx], Z[(n - kx) % n], xs[q][0], xs[q][1]);
            }
            __syncthreads();
            #pragma unroll
            for (int q = 0; q < MAXQ; ++q)
            {
                const int kx = lane + 64*q;
                if (kx < nx) stage_store<TF>(A, kx, n, xs[q][0], xs[q][1]);
            }
            soff = 0;
        }
        else
        {
            soff = (Z == A) ? n : 0;
            T* S = rows + row*re + soff;
            for (int kx = lane; kx < nx; kx += 64)
            {
                T xa, xb;
                unpack_pair<TF>(Z[kx], Z[(n - kx) % n], xa, xb);
                stage_store<TF>(S, kx, n, xa, xb);
            }
        }
        __syncthreads();
        // The block writes C[k][kx][j0 .. j0+2RB) with consecutive threads on consecutive j.
        const int nthr = 64*blockDim.y, tid = lane + 64*row;
        const int JB = 2*RB;
        T* Ck = C + static_cast<size_t>(k)*nx*M;
        for (int idx = tid; idx < nx*JB; idx += nthr)
        {
            const int kx = idx / JB, jl = idx % JB;
            const int r = jl >> 1, part = jl & 1;
            const int j = j0 + jl;
            if (j >= M) continue;
            const T* Sr = rows + r*re + soff;
            T val;
            if (kx == 0 || 2*kx == n)
            {
                const T d = Sr[kx == 0 ? 0 : 1];
                val.x = part ? d.y : d.x; val.y = TF(0);
            }
            else
                val = Sr[stage_base(kx, n) + part];
            Ck[static_cast<size_t>(kx)*M + j] = val;
        }
    }

    // y transform (forward sign -1 or inverse +1) of RB contiguous rows of C;
    // the inverse divides by jtot as src/fft.cxx:420 does.
    template<typename TF>
    __global__ void __launch_bounds__(256) fft_y_g(cplx<TF>* __restrict__ C, Plan plan, const double2* __restrict__ tw,
                                                   int sign, int nrows, int RB)
    {
        using T = cplx<TF>;
        extern __shared__ unsigned char smem_raw[];
        T* tws = reinterpret_cast<T*>(smem_raw);
        const int n = plan.n, re = row_elems(plan);
        const int row = threadIdx.y, lane = threadIdx.x;
        const int r = blockIdx.x*RB + row;
        T* A = tws + n + row*re;
        if (plan.fast) load_twiddles<TF>(tws, tw, n);
        const bool active = r < nrows;
        T* Cr = C + static_cast<size_t>(active ? r : 0)*n;
        if (active)
            for (int x = lane; x < n; x += 64) A[x] = Cr[x];
        __syncthreads();
        T* Z = fft_row<TF>(A, tws, tw, plan, sign, lane);
        if (!active) return;
        if (sign < 0)
            for (int x = lane; x < n; x += 64) Cr[x] = Z[x];
        else
            for (int x = lane; x < n; x += 64) { T z = Z[x]; z.x /= TF(n); z.y /= TF(n); Cr[x] = z; }
    }

    // Tridiagonal solve per (kx, ky) mode (Pres_2::solve + tdma,
    // src/pres_2.cxx:266-324, 200-263), one thread per mode marching in k.
    template<typename TF>
    __global__ void __launch_bounds__(256) tdma_g(cplx<TF>* __restrict__ C, TF* __restrict__ gam,
                                                  const TF* __restrict__ bmati, const TF* __restrict__ bmatj,
                                                  const TF* __restrict__ a, const TF* __restrict__ c,
                                                  const TF* __restrict__ dz, const TF* __restrict__ rho,
                                                  int nx, Gd g)
    {
        using T = cplx<TF>;
        const int M = g.jtot, K = g.ktot, kgc = g.kgc;
        const int mode = blockIdx.x*blockDim.x + threadIdx.x;
        const int nmodes = nx*M;
        if (mode >= nmodes) return;
        const int kx = mode / M, ky = mode % M;
        const TF lam = bmati[kx] + bmatj[ky];
        const size_t kstride = static_cast<size_t>(nmodes);

        auto bcoef = [&](int k) -> TF {
            TF b = dz[k+kgc]*dz[k+kgc] * rho[k+kgc]*lam - (a[k] + c[k]);
            if (k == 0) b += a[0];
            if (k == K-1) { if (kx == 0 && ky == 0) b -= c[K-1]; else b += c[K-1]; }
            return b;
        };

        // forward elimination
        TF work2d = bcoef(0);
        T p = C[mode];
        TF f = dz[kgc]*dz[kgc];
        p.x = f*p.x; p.y = f*p.y;
        p.x /= work2d; p.y /= work2d;
        C[mode] = p;
        T pprev = p;
        for (int k = 1; k < K; ++k)
        {
            const TF gk = c[k-1] / work2d;
            gam[k*kstride + mode] = gk;
            work2d = bcoef(k) - a[k]*gk;
            T q = C[k*kstride + mode];
            const TF fk = dz[k+kgc]*dz[k+kgc];
            q.x = fk*q.x; q.y = fk*q.y;
            q.x -= a[k]*pprev.x; q.y -= a[k]*pprev.y;
            q.x /= work2d; q.y /= work2d;
            C[k*kstride + mode] = q;
            pprev = q;
        }
        // back substitution
        T pn = pprev;
        for (int k = K-2; k >= 0; --k)
        {
            T q = C[k*kstride + mode];
            const TF gk1 = gam[(k+1)*kstride + mode];
            q.x -= gk1*pn.x; q.y -= gk1*pn.y;
            C[k*kstride + mode] = q;
            pn = q;
        }
    }

    // Inverse x transform and scatter into the ghosted p (Pres_2::solve tail,
    // src/pres_2.cxx:336-361): interior, the periodic halos of the rows the
    // block owns, and the bottom ghost level p[kstart-1] = p[kstart].
    template<typename TF>
    __global__ void __launch_bounds__(256) fft_x_bwd_g(const cplx<TF>* __restrict__ C, TF* __restrict__ p,
                                                       Plan plan, const double2* __restrict__ tw, int RB, Gd g)
    {
        using T = cplx<TF>;
        extern __shared__ unsigned char smem_raw[];
        T* tws = reinterpret_cast<T*>(smem_raw);
        const int n = plan.n, nx = n/2 + 1, M = g.jtot, re = row_elems(plan);
        T* rows = tws + n;
        const int row = threadIdx.y, lane = threadIdx.x;
        const int k = blockIdx.y;
        const int j0 = blockIdx.x*2*RB;
        const int nthr = 64*blockDim.y, tid = lane + 64*row;
        const int JB = 2*RB;
        if (plan.fast) load_twiddles<TF>(tws, tw, n);
        const T* Ck = C + static_cast<size_t>(k)*nx*M;
        // Stage the block's spectra packed as in the forward kernel; the DC and
        // Nyquist entries of a real row are real.  The staging buffer is the row
        // buffer itself (in-place engine) or the row's second buffer (fallback).
        const int soff = plan.fast ? 0 : n;
        for (int idx = tid; idx < nx*JB; idx += nthr)
        {
            const int kx = idx / JB, jl = idx % JB;
            const int r = jl >> 1, part = jl & 1;
            const int j = j0 + jl;
            T val;
            if (j < M) val = Ck[static_cast<size_t>(kx)*M + j];
            else { val.x = TF(0); val.y = TF(0); }
            T* Sr = rows + r*re + soff;
            if (kx == 0 || 2*kx == n)
            {
                TF* d = reinterpret_cast<TF*>(&Sr[kx == 0 ? 0 : 1]);
                d[part] = val.x;
            }
            else
                Sr[stage_base(kx, n) + part] = val;
        }
        __syncthreads();
        T* A = rows + row*re;
        const T* B = A + soff;
        // Hermitian extension of both spectra and packing Z = Xa + i Xb.
        auto zval = [&](int x) -> T {
            T xa, xb;
            const int kx = (x <= n/2) ? x : n - x;
            if (kx == 0 || 2*kx == n)
            {
                const T d = B[kx == 0 ? 0 : 1];
                xa.x = d.x; xa.y = TF(0); xb.x = d.y; xb.y = TF(0);
            }
            else
            {
                const int b = stage_base(kx, n);
                xa = B[b]; xb = B[b+1];
                if (x > n/2) { xa.y = -xa.y; xb.y = -xb.y; }
            }
            T z; z.x = xa.x - xb.y; z.y = xa.y + xb.x;
            return z;
        };
        if (plan.fast)
        {
            constexpr int MAXX = 2*MAXQ;   // n/64 values per lane, n <= 1024
            T zs[MAXX];
            #pragma unroll
            for (int q = 0; q < MAXX; ++q)
            {
                const int x = lane + 64*q;
                if (x < n) zs[q] = zval(x);
            }
            __syncthreads();
            #pragma unroll
            for (int q = 0; q < MAXX; ++q)
            {
                const int x = lane + 64*q;
                if (x < n) A[x] = zs[q];
            }
        }
        else
            for (int x = lane; x < n; x += 64) A[x] = zval(x);
        __syncthreads();
        T* Z = fft_row<TF>(A, tws, tw, plan, +1, lane);
        const int ja = j0 + 2*row, jb = ja + 1;
        const int jj = g.icells, kk = g.ijcells;
        const int kc = k + g.kgc;
        for (int pr = 0; pr < 2; ++pr)
        {
            const int j = pr ? jb : ja;
            if (j >= M) continue;
            // destination rows: the interior row and, within jgc of an edge, its periodic image
            int rows_[3]; int nrow = 0;
            rows_[nrow++] = j + g.jgc;
            if (j >= M - g.jgc) rows_[nrow++] = j - (M - g.jgc);
            if (j < g.jgc) rows_[nrow++] = g.jend + j;
            for (int x = lane; x < g.icells; x += 64)
            {
                // ghosted column x maps to interior column x - igc, wrapped
                int xi = x - g.igc;
                xi = (xi < 0) ? xi + n : (xi >= n ? xi - n : xi);
                const T z = Z[xi];
                const TF val = (pr ? z.y : z.x) / TF(n);
                for (int q = 0; q < nrow; ++q)
                {
                    const int ij = x + rows_[q]*jj;
                    p[ij + kc*kk] = val;
                    if (k == 0) p[ij + (kc-1)*kk] = val;
                }
            }
        }
    }

    // Pres_2::output (src/pres_2.cxx:364-387).
    template<typename TF>
    __global__ void __launch_bounds__(BX*BY) pres_out_g(TF* __restrict__ ut, TF* __restrict__ vt, TF* __restrict__ wt,
                                                        const TF* __restrict__ p, const TF* __restrict__ dzhi,
                                                        TF dxi, TF dyi, Gd g)
    {
        const int i = blockIdx.x*BX + threadIdx.x + g.istart;
        const int j = blockIdx.y*BY + threadIdx.y + g.jstart;
        const int k = blockIdx.z + g.kstart;
        if (i >= g.iend || j >= g.jend) return;
        const int ijk = i + j*g.icells + k*g.ijcells;
        const TF pc = p[ijk];
        ut[ijk] -= (pc - p[ijk-1]) * dxi;
        vt[ijk] -= (pc - p[ijk-g.icells]) * dyi;
        wt[ijk] -= (pc - p[ijk-g.ijcells]) * dzhi[k];
    }

    // Pres_2::output followed by the RK3 update of u, v, w (src/pres_2.cxx:364-387,
    // src/timeloop.cxx:251-286): the same operations in one pass.
    template<typename TF>
    __global__ void __launch_bounds__(BX*BY) pres_out_rk3_g(
            TF* __restrict__ u, TF* __restrict__ v, TF* __restrict__ w,
            TF* __restrict__ ut, TF* __restrict__ vt, TF* __restrict__ wt,
            const TF* __restrict__ p, const TF* __restrict__ dzhi, TF dxi, TF dyi,
            TF cBdt, TF cAn, int zero, Gd g)
    {
        const int i = blockIdx.x*BX + threadIdx.x + g.istart;
        const int j = blockIdx.y*BY + threadIdx.y + g.jstart;
        const int k = blockIdx.z + g.kstart;
        if (i >= g.iend || j >= g.jend) return;
        const int ijk = i + j*g.icells + k*g.ijcells;
        const TF pc = p[ijk];
        TF tu = ut[ijk]; tu -= (pc - p[ijk-1]) * dxi;
        TF tv = vt[ijk]; tv -= (pc - p[ijk-g.icells]) * dyi;
        TF tw = wt[ijk]; tw -= (pc - p[ijk-g.ijcells]) * dzhi[k];
        u[ijk] += cBdt*tu; ut[ijk] = zero ? TF(0) : tu*cAn;
        v[ijk] += cBdt*tv; vt[ijk] = zero ? TF(0) : tv*cAn;
        w[ijk] += cBdt*tw; wt[ijk] = zero ? TF(0) : tw*cAn;
    }

    // Pres_2::calc_divergence (src/pres_2.cxx:389-422), per-block maxima.
    template<typename TF>
    __global__ void __launch_bounds__(BX*BY) divergence_g(const TF* __restrict__ u, const TF* __restrict__ v,
                                                          const TF* __restrict__ w, const TF* __restrict__ dzi,
                                                          const TF* __restrict__ rho, const TF* __restrict__ rhoh,
                                                          TF dxi, TF dyi, Gd g, double* __restrict__ red)
    {
        __shared__ double sm[BX*BY];
        const int tid = threadIdx.x + threadIdx.y*BX;
        const int i = blockIdx.x*BX + threadIdx.x + g.istart;
        const int j = blockIdx.y*BY + threadIdx.y + g.jstart;
        const int kb = g.kstart + blockIdx.z*RED_KCHUNK;
        const int ke = min(kb + RED_KCHUNK, g.kend);
        double m = 0.;
        if (i < g.iend && j < g.jend)
            for (int k = kb; k < ke; ++k)
            {
                const int jj = g.icells, kk = g.ijcells;
                const int ijk = i + j*jj + k*kk;
                const TF div = rho[k]*((u[ijk+1]-u[ijk])*dxi + (v[ijk+jj]-v[ijk])*dyi)
                             + (rhoh[k+1]*w[ijk+kk]-rhoh[k]*w[ijk])*dzi[k];
                m = fmax(m, static_cast<double>(fabs(div)));
            }
        sm[tid] = m;
        __syncthreads();
        for (int s = BX*BY/2; s > 0; s >>= 1)
        {
            if (tid < s) sm[tid] = fmax(sm[tid], sm[tid + s]);
            __syncthreads();
        }
        if (tid == 0)
            red[1 + blockIdx.x + gridDim.x*(blockIdx.y + gridDim.y*blockIdx.z)] = sm[0];
    }

    // Rows per block so that the twiddle table and the row buffers fit in 64 KiB.
    int rows_per_block(const Plan& p, size_t elem)
    {
        const size_t budget = 65536 / elem;          // complex elements
        const size_t per_row = static_cast<size_t>(row_elems(p));
        const size_t avail = budget > static_cast<size_t>(p.n) ? budget - p.n : 0;
        return std::max(1, std::min(4, static_cast<int>(avail / per_row)));
    }

    size_t lds_bytes(const Plan& p, int RB, size_t elem)
    {
        return (static_cast<size_t>(p.n) + static_cast<size_t>(RB)*row_elems(p))*elem;
    }

    template<typename TF>
    int setup_tables(mhh_ctx* c, Pres_tables<TF>& t)
    {
        const Gd& g = c->gd;
        const int N = g.itot, M = g.jtot, K = g.ktot;
        // Pres_2::set_values (src/pres_2.cxx:124-153) in the element type.
        std::vector<TF> bi(N), bj(M), av(K), cv(K);
        const TF dxidxi = TF(1.)/(TF(g.dx)*TF(g.dx));
        const TF dyidyi = TF(1.)/(TF(g.dy)*TF(g.dy));
        const TF pi = std::acos(TF(-1.));
        for (int j = 0; j < M/2+1; ++j) bj[j] = TF(2.) * (std::cos(TF(2.)*pi*TF(j)/TF(M)) - TF(1.)) * dyidyi;
        for (int j = M/2+1; j < M; ++j) bj[j] = bj[M-j];
        for (int i = 0; i < N/2+1; ++i) bi[i] = TF(2.) * (std::cos(TF(2.)*pi*TF(i)/TF(N)) - TF(1.)) * dxidxi;
        for (int i = N/2+1; i < N; ++i) bi[i] = bi[N-i];
        for (int k = 0; k < K; ++k)
        {
            av[k] = TF(c->h_dz[k+g.kgc]) * TF(c->h_rhorefh[k+g.kgc  ])*TF(c->h_dzhi[k+g.kgc  ]);
            cv[k] = TF(c->h_dz[k+g.kgc]) * TF(c->h_rhorefh[k+g.kgc+1])*TF(c->h_dzhi[k+g.kgc+1]);
        }
        auto up = [](TF** d, const std::vector<TF>& h) -> int {
            MHH_CHECK(check_hip(hipMalloc(d, sizeof(TF)*h.size()), "hipMalloc(pres table)"));
            return check_hip(hipMemcpy(*d, h.data(), sizeof(TF)*h.size(), hipMemcpyHostToDevice), "hipMemcpy(pres table)");
        };
        MHH_CHECK(up(&t.bmati, bi));
        MHH_CHECK(up(&t.bmatj, bj));
        MHH_CHECK(up(&t.a, av));
        MHH_CHECK(up(&t.c, cv));
        // Twiddles exp(-2 pi i t/n) in long double, rounded once.
        auto roots = [](double** d, int n) -> int {
            std::vector<double> h(2*static_cast<size_t>(n));
            const long double pil = 3.141592653589793238462643383279502884L;
            for (int t = 0; t < n; ++t)
            {
                const long double ang = 2.0L*pil*static_cast<long double>(t)/static_cast<long double>(n);
                h[2*t] = static_cast<double>(cosl(ang));
                h[2*t+1] = static_cast<double>(-sinl(ang));
            }
            MHH_CHECK(check_hip(hipMalloc(d, sizeof(double)*h.size()), "hipMalloc(twiddles)"));
            return check_hip(hipMemcpy(*d, h.data(), sizeof(double)*h.size(), hipMemcpyHostToDevice), "hipMemcpy(twiddles)");
        };
        MHH_CHECK(roots(&t.twx, N));
        MHH_CHECK(roots(&t.twy, M));
        return 0;
    }

    template<typename TF>
    void free_tables(Pres_tables<TF>& t)
    {
        if (t.bmati) (void)hipFree(t.bmati);
        if (t.bmatj) (void)hipFree(t.bmatj);
        if (t.a) (void)hipFree(t.a);
        if (t.c) (void)hipFree(t.c);
        if (t.twx) (void)hipFree(t.twx);
        if (t.twy) (void)hipFree(t.twy);
        t = Pres_tables<TF>();
    }

    // rk3 >= 0: apply the gradient together with the RK3 update of substep
    // rk3 (time step rkdt) instead of Pres_2::output alone.
    template<typename TF>
    int exec(mhh_ctx* c, const Prof<TF>& p, const Pres_tables<TF>& t, double dt, int rk3 = -1, double rkdt = 0.)
    {
        using T = cplx<TF>;
        const Gd& g = c->gd;
        if (!c->f.p || !c->f.tmp1 || !c->f.tmp2)
            return fail(MHH_ESTATE, "pres_2 needs p, tmp1 and tmp2 bound");
        const int N = g.itot, M = g.jtot, K = g.ktot, nx = N/2 + 1;
        Plan px, py;
        if (!make_plan(N, px) || !make_plan(M, py))
            return fail(MHH_EINVAL, "pres_2: itot and jtot must factor into primes <= 13");

        TF* P = static_cast<TF*>(c->f.p);
        T* C = static_cast<T*>(c->f.tmp1);
        TF* gam = static_cast<TF*>(c->f.tmp2);
        const TF dxi = TF(1.)/TF(g.dx), dyi = TF(1.)/TF(g.dy), dti = TF(1.)/TF(dt);

        // 1. divergence into the dense array P (stored in p)
        {
            const dim3 grd((N + BX - 1)/BX, (M + BY - 1)/BY, K);
            { mhh::Timer _t(c, "pres_in"); pres_in_g<TF><<<grd, dim3(BX, BY), 0, c->stream>>>(P,
                    static_cast<const TF*>(c->f.u), static_cast<const TF*>(c->f.v), static_cast<const TF*>(c->f.w),
                    static_cast<const TF*>(c->f.ut), static_cast<const TF*>(c->f.vt), static_cast<const TF*>(c->f.wt),
                    p.dzi, p.rhoref, p.rhorefh, dxi, dyi, dti, g); }
            MHH_CHECK(check_launch("pres_in"));
        }
        // 2. forward x transform
        {
            const int RB = rows_per_block(px, sizeof(T));
            const dim3 grd((M + 2*RB - 1)/(2*RB), K);
            const size_t lds = lds_bytes(px, RB, sizeof(T));
            { mhh::Timer _t(c, "fft_x_fwd"); fft_x_fwd_g<TF><<<grd, dim3(64, RB), lds, c->stream>>>(P, C, px, reinterpret_cast<const double2*>(t.twx), RB, g); }
            MHH_CHECK(check_launch("fft_x_fwd"));
        }
        // 3. forward y transform
        const int nrows = nx*K;
        const int RBy = rows_per_block(py, sizeof(T));
        const size_t ldsy = lds_bytes(py, RBy, sizeof(T));
        { mhh::Timer _t(c, "fft_y_fwd"); fft_y_g<TF><<<(nrows + RBy - 1)/RBy, dim3(64, RBy), ldsy, c->stream>>>(C, py,
                reinterpret_cast<const double2*>(t.twy), -1, nrows, RBy); }
        MHH_CHECK(check_launch("fft_y_fwd"));
        // 4. tridiagonal per mode
        {
            const int nmodes = nx*M;
            { mhh::Timer _t(c, "tdma"); tdma_g<TF><<<(nmodes + 255)/256, 256, 0, c->stream>>>(C, gam, t.bmati, t.bmatj, t.a, t.c,
                                                                  p.dz, p.rhoref, nx, g); }
            MHH_CHECK(check_launch("tdma"));
        }
        // 5. inverse y transform (divides by jtot)
        { mhh::Timer _t(c, "fft_y_bwd"); fft_y_g<TF><<<(nrows + RBy - 1)/RBy, dim3(64, RBy), ldsy, c->stream>>>(C, py,
                reinterpret_cast<const double2*>(t.twy), +1, nrows, RBy); }
        MHH_CHECK(check_launch("fft_y_bwd"));
        // 6. inverse x transform (divides by itot) into the ghosted p with halos
        {
            const int RB = rows_per_block(px, sizeof(T));
            const dim3 grd((M + 2*RB - 1)/(2*RB), K);
            const size_t lds = lds_bytes(px, RB, sizeof(T));
            { mhh::Timer _t(c, "fft_x_bwd"); fft_x_bwd_g<TF><<<grd, dim3(64, RB), lds, c->stream>>>(C, P, px, reinterpret_cast<const double2*>(t.twx), RB, g); }
            MHH_CHECK(check_launch("fft_x_bwd"));
        }
        // 7. pressure gradient on the tendencies
        if (rk3 < 0)
        {
            { mhh::Timer _t(c, "pres_out"); pres_out_g<TF><<<grid3(g, g.kstart, g.kend), dim3(BX, BY), 0, c->stream>>>(
                    static_cast<TF*>(c->f.ut), static_cast<TF*>(c->f.vt), static_cast<TF*>(c->f.wt),
                    P, p.dzhi, dxi, dyi, g); }
            return check_launch("pres_out");
        }
        // Williamson low-storage coefficients (src/timeloop.cxx:255-256).
        const TF cA[] = {TF(0.), TF(-5./9.), TF(-153./128.)};
        const TF cB[] = {TF(1./3.), TF(15./16.), TF(8./15.)};
        const int substepn = (rk3 + 1) % 3;
        { mhh::Timer _t(c, "pres_out_rk3"); pres_out_rk3_g<TF><<<grid3(g, g.kstart, g.kend), dim3(BX, BY), 0, c->stream>>>(
                static_cast<TF*>(c->f.u), static_cast<TF*>(c->f.v), static_cast<TF*>(c->f.w),
                static_cast<TF*>(c->f.ut), static_cast<TF*>(c->f.vt), static_cast<TF*>(c->f.wt),
                P, p.dzhi, dxi, dyi, cB[rk3]*TF(rkdt), cA[substepn], substepn == 0, g); }
        MHH_CHECK(check_launch("pres_out_rk3"));
        if (substepn == 0)
        {
            MHH_CHECK(zero_tendency_ghosts<TF>(c, static_cast<TF*>(c->f.ut)));
            MHH_CHECK(zero_tendency_ghosts<TF>(c, static_cast<TF*>(c->f.vt)));
            MHH_CHECK(zero_tendency_ghosts<TF>(c, static_cast<TF*>(c->f.wt)));
        }
        return 0;
    }

    template<typename TF>
    int divergence(mhh_ctx* c, const Prof<TF>& p, double* out)
    {
        const Gd& g = c->gd;
        const dim3 grd = grid_red(g);
        const int nb = static_cast<int>(grd.x*grd.y*grd.z);
        MHH_CHECK(ensure_reduction(c, nb));
        { mhh::Timer _t(c, "divergence"); divergence_g<TF><<<grd, dim3(BX, BY), 0, c->stream>>>(
                static_cast<const TF*>(c->f.u), static_cast<const TF*>(c->f.v), static_cast<const TF*>(c->f.w),
                p.dzi, p.rhoref, p.rhorefh, TF(1.)/TF(g.dx), TF(1.)/TF(g.dy), g, c->red); }
        MHH_CHECK(check_launch("divergence"));
        return reduce_max_partials(c, nb, out);
    }
}

namespace mhh
{
    int pres_init(mhh_ctx* c)
    {
        Plan px, py;
        if (!make_plan(c->gd.itot, px) || !make_plan(c->gd.jtot, py))
            return fail(MHH_EINVAL, "itot and jtot must factor into primes <= 13 (FFT radices)");
        return c->precision == MHH_FP64 ? setup_tables<double>(c, c->td) : setup_tables<float>(c, c->tf);
    }

    void pres_free(mhh_ctx* c)
    {
        free_tables(c->td);
        free_tables(c->tf);
    }
}

extern "C" {

int mhh_pres_2_exec(mhh_ctx* c, double dt)
{
    MHH_REQUIRE_BOUND(c);
    if (!(dt > 0)) return mhh::fail(MHH_EINVAL, "pres_2: dt must be positive");
    return c->precision == MHH_FP64 ? exec<double>(c, c->pd, c->td, dt) : exec<float>(c, c->pf, c->tf, dt);
}

int mhh_pres_2_exec_rk3(mhh_ctx* c, double dt, int substep, double rkdt)
{
    MHH_REQUIRE_BOUND(c);
    if (!(dt > 0)) return mhh::fail(MHH_EINVAL, "pres_2: dt must be positive");
    if (substep < 0 || substep > 2) return mhh::fail(MHH_EINVAL, "rk3 substep must be 0, 1 or 2");
    return c->precision == MHH_FP64 ? exec<double>(c, c->pd, c->td, dt, substep, rkdt)
                                    : exec<float>(c, c->pf, c->tf, dt, substep, rkdt);
}

int mhh_pres_2_divergence(mhh_ctx* c, double* out)
{
    MHH_REQUIRE_BOUND(c);
    if (!out) return mhh::fail(MHH_EINVAL, "null output");
    return c->precision == MHH_FP64 ? divergence<double>(c, c->pd, out) : divergence<float>(c, c->pf, out);
}

} // extern "C"
