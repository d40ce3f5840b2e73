/*
 * ps_oracle.c -- CPU restatement of the reference's Threads backend for the
 * test-bed kernels. TEST INFRASTRUCTURE ONLY: used by tests/ as the checker
 * and by bench.py as the timed CPU baseline; never part of the product path.
 *
 * Semantics restated (ParallelStencil.jl v0.16.0):
 *   - loop nest: `Threads.@threads for iz; for iy; for ix; body` over the
 *     ranges 1:maxsize of all array arguments
 *     (src/ParallelKernel/parallel.jl:715-768, 785-798) -> OpenMP static
 *     schedule over the outermost dimension;
 *   - every statement guarded by its @within mask (src/parallel.jl:654-680,
 *     src/FiniteDifferences.jl:397-407);
 *   - on-the-fly arrays recomputed at each shifted use (src/parallel.jl:789-859);
 *   - index roles ixi = ix+1, ixd = ix (padding=false,
 *     src/ParallelKernel/parallel.jl:498-504);
 *   - Julia arithmetic: n-ary + and * fold left, unary minus binds tighter
 *     than *, no FMA contraction (build with -ffp-contract=off), IEEE division.
 * Arrays are dense column-major (x fastest); indices below are 1-based.
 *
 * Parity pinning of every function: see the header of ps_oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define IX3(i, j, k, nx, ny) ((size_t)((i)-1) + (size_t)((j)-1) * (size_t)(nx) + (size_t)((k)-1) * (size_t)(nx) * (size_t)(ny))
#define IX2(i, j, nx) ((size_t)((i)-1) + (size_t)((j)-1) * (size_t)(nx))

/* ---------------------------------------------------------------- BFloat16 */
static inline float bf16_ld(uint16_t v) {
  uint32_t u = (uint32_t)v << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}
static inline uint16_t bf16_st(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0u;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

/* ------------------------------------------------------------------ triad */
/* examples/memcopy3D.jl:10-13: @all(T2) = @all(T) + @all(Ci). */
void pso_triad(int64_t n, double* T2, const double* T, const double* Ci) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) T2[i] = T[i] + Ci[i];
}

/* ------------------------------------------------------ diffusion (flux) */
/* Flux form of the 3-D heat diffusion step with on-the-fly fluxes:
 *   @all(qTx) = -lam*@d_xi(T)/dx   (likewise y, z)
 *   @inn(T2)  = @inn(T) + dt/@inn(Cp)*(-@d_xa(qTx)/dx - @d_ya(qTy)/dy - @d_za(qTz)/dz)
 * (test/test_parallel.jl:419-426 with the division form of the configuration). */
void pso_diffusion3d_flux(int nx, int ny, int nz, double* T2, const double* T, const double* Cp, double lam, double dt,
                          double dx, double dy, double dz) {
#define T_(i, j, k) T[IX3(i, j, k, nx, ny)]
#pragma omp parallel for schedule(static)
  for (int iz = 1; iz <= nz; ++iz)
    for (int iy = 1; iy <= ny; ++iy)
      for (int ix = 1; ix <= nx; ++ix) {
        if (!(1 <= ix && ix <= nx - 2 && 1 <= iy && iy <= ny - 2 && 1 <= iz && iz <= nz - 2)) continue;
        /* qTx(i,j,k) = ((-lam)*(T[i+1,j+1,k+1] - T[i,j+1,k+1]))/dx */
        double qx1 = ((-lam) * (T_(ix + 2, iy + 1, iz + 1) - T_(ix + 1, iy + 1, iz + 1))) / dx;
        double qx0 = ((-lam) * (T_(ix + 1, iy + 1, iz + 1) - T_(ix, iy + 1, iz + 1))) / dx;
        double qy1 = ((-lam) * (T_(ix + 1, iy + 2, iz + 1) - T_(ix + 1, iy + 1, iz + 1))) / dy;
        double qy0 = ((-lam) * (T_(ix + 1, iy + 1, iz + 1) - T_(ix + 1, iy, iz + 1))) / dy;
        double qz1 = ((-lam) * (T_(ix + 1, iy + 1, iz + 2) - T_(ix + 1, iy + 1, iz + 1))) / dz;
        double qz0 = ((-lam) * (T_(ix + 1, iy + 1, iz + 1) - T_(ix + 1, iy + 1, iz))) / dz;
        T2[IX3(ix + 1, iy + 1, iz + 1, nx, ny)] =
            T_(ix + 1, iy + 1, iz + 1) +
            (dt / Cp[IX3(ix + 1, iy + 1, iz + 1, nx, ny)]) * ((((-(qx1 - qx0)) / dx) - ((qy1 - qy0) / dy)) - ((qz1 - qz0) / dz));
      }
#undef T_
}

/* The on-the-fly kernel of test/test_parallel.jl:419-426 verbatim in form:
 *   @all(qx) = -lam*@d_xi(T)*_dx ; @all(dTdt) = @inn(Ci)*(-@d_xa(qx)*_dx - @d_ya(qy)*_dy - @d_za(qz)*_dz)
 *   @inn(T2) = @inn(T) + dt*@all(dTdt) */
void pso_diffusion3d_otf_test(int nx, int ny, int nz, double* T2, const double* T, const double* Ci, double lam, double dt,
                              double _dx, double _dy, double _dz) {
#define T_(i, j, k) T[IX3(i, j, k, nx, ny)]
#pragma omp parallel for schedule(static)
  for (int iz = 1; iz <= nz; ++iz)
    for (int iy = 1; iy <= ny; ++iy)
      for (int ix = 1; ix <= nx; ++ix) {
        if (!(1 <= ix && ix <= nx - 2 && 1 <= iy && iy <= ny - 2 && 1 <= iz && iz <= nz - 2)) continue;
        double qx1 = ((-lam) * (T_(ix + 2, iy + 1, iz + 1) - T_(ix + 1, iy + 1, iz + 1))) * _dx;
        double qx0 = ((-lam) * (T_(ix + 1, iy + 1, iz + 1) - T_(ix, iy + 1, iz + 1))) * _dx;
        double qy1 = ((-lam) * (T_(ix + 1, iy + 2, iz + 1) - T_(ix + 1, iy + 1, iz + 1))) * _dy;
        double qy0 = ((-lam) * (T_(ix + 1, iy + 1, iz + 1) - T_(ix + 1, iy, iz + 1))) * _dy;
        double qz1 = ((-lam) * (T_(ix + 1, iy + 1, iz + 2) - T_(ix + 1, iy + 1, iz + 1))) * _dz;
        double qz0 = ((-lam) * (T_(ix + 1, iy + 1, iz + 1) - T_(ix + 1, iy + 1, iz))) * _dz;
        double dTdt = Ci[IX3(ix + 1, iy + 1, iz + 1, nx, ny)] * ((((-(qx1 - qx0)) * _dx) - ((qy1 - qy0) * _dy)) - ((qz1 - qz0) * _dz));
        T2[IX3(ix + 1, iy + 1, iz + 1, nx, ny)] = T_(ix + 1, iy + 1, iz + 1) + dt * dTdt;
      }
#undef T_
}

/* examples/diffusion3D_novis_noperf.jl:10-13:
 *   @inn(T2) = @inn(T) + dt*(lam*@inn(Ci)*(@d2_xi(T)/dx^2 + @d2_yi(T)/dy^2 + @d2_zi(T)/dz^2)) */
void pso_diffusion3d_d2(int nx, int ny, int nz, double* T2, const double* T, const double* Ci, double lam, double dt, double dx,
                        double dy, double dz) {
#define T_(i, j, k) T[IX3(i, j, k, nx, ny)]
#pragma omp parallel for schedule(static)
  for (int iz = 1; iz <= nz; ++iz)
    for (int iy = 1; iy <= ny; ++iy)
      for (int ix = 1; ix <= nx; ++ix) {
        if (!(1 <= ix && ix <= nx - 2 && 1 <= iy && iy <= ny - 2 && 1 <= iz && iz <= nz - 2)) continue;
        double c = T_(ix + 1, iy + 1, iz + 1);
        double d2x = (T_(ix + 2, iy + 1, iz + 1) - c) - (c - T_(ix, iy + 1, iz + 1));
        double d2y = (T_(ix + 1, iy + 2, iz + 1) - c) - (c - T_(ix + 1, iy, iz + 1));
        double d2z = (T_(ix + 1, iy + 1, iz + 2) - c) - (c - T_(ix + 1, iy + 1, iz));
        T2[IX3(ix + 1, iy + 1, iz + 1, nx, ny)] =
            c + dt * ((lam * Ci[IX3(ix + 1, iy + 1, iz + 1, nx, ny)]) * (((d2x / (dx * dx)) + (d2y / (dy * dy))) + (d2z / (dz * dz))));
      }
#undef T_
}

/* ------------------------------------------------------------ acoustic 3-D */
/* miniapps/acoustic3D.jl:11-16 compute_V!, ranges 1:(n+1) in every dimension.
 * P: nx*ny*nz, Vx: (nx+1)*ny*nz, Vy: nx*(ny+1)*nz, Vz: nx*ny*(nz+1). */
#define ACOUSTIC3D_V_BODY(LD, ST, T)                                                                                          \
  for (int iz = 1; iz <= nz + 1; ++iz)                                                                                      \
    for (int iy = 1; iy <= ny + 1; ++iy)                                                                                    \
      for (int ix = 1; ix <= nx + 1; ++ix) {                                                                                \
        if (1 <= ix && ix <= nx - 1 && 1 <= iy && iy <= ny - 2 && 1 <= iz && iz <= nz - 2) {                                \
          size_t v = IX3(ix + 1, iy + 1, iz + 1, nx + 1, ny);                                                               \
          T dP = LD(P[IX3(ix + 1, iy + 1, iz + 1, nx, ny)]) - LD(P[IX3(ix, iy + 1, iz + 1, nx, ny)]);                        \
          Vx[v] = ST(LD(Vx[v]) - ((dt / rho) * dP) / dx);                                                                   \
        }                                                                                                                   \
        if (1 <= ix && ix <= nx - 2 && 1 <= iy && iy <= ny - 1 && 1 <= iz && iz <= nz - 2) {                                \
          size_t v = IX3(ix + 1, iy + 1, iz + 1, nx, ny + 1);                                                               \
          T dP = LD(P[IX3(ix + 1, iy + 1, iz + 1, nx, ny)]) - LD(P[IX3(ix + 1, iy, iz + 1, nx, ny)]);                        \
          Vy[v] = ST(LD(Vy[v]) - ((dt / rho) * dP) / dy);                                                                   \
        }                                                                                                                   \
        if (1 <= ix && ix <= nx - 2 && 1 <= iy && iy <= ny - 2 && 1 <= iz && iz <= nz - 1) {                                \
          size_t v = IX3(ix + 1, iy + 1, iz + 1, nx, ny);                                                                   \
          T dP = LD(P[IX3(ix + 1, iy + 1, iz + 1, nx, ny)]) - LD(P[IX3(ix + 1, iy + 1, iz, nx, ny)]);                        \
          Vz[v] = ST(LD(Vz[v]) - ((dt / rho) * dP) / dz);                                                                   \
        }                                                                                                                   \
      }

/* miniapps/acoustic3D.jl:18-21 compute_P!:
 *   @all(P) = @all(P) - dt*k*(@d_xa(Vx)/dx + @d_ya(Vy)/dy + @d_za(Vz)/dz) */
#define ACOUSTIC3D_P_BODY(LD, ST, T)                                                                                          \
  for (int iz = 1; iz <= nz + 1; ++iz)                                                                                      \
    for (int iy = 1; iy <= ny + 1; ++iy)                                                                                    \
      for (int ix = 1; ix <= nx + 1; ++ix) {                                                                                \
        if (!(1 <= ix && ix <= nx && 1 <= iy && iy <= ny && 1 <= iz && iz <= nz)) continue;                                 \
        T dvx = LD(Vx[IX3(ix + 1, iy, iz, nx + 1, ny)]) - LD(Vx[IX3(ix, iy, iz, nx + 1, ny)]);                              \
        T dvy = LD(Vy[IX3(ix, iy + 1, iz, nx, ny + 1)]) - LD(Vy[IX3(ix, iy, iz, nx, ny + 1)]);                              \
        T dvz = LD(Vz[IX3(ix, iy, iz + 1, nx, ny)]) - LD(Vz[IX3(ix, iy, iz, nx, ny)]);                                      \
        size_t p = IX3(ix, iy, iz, nx, ny);                                                                                 \
        P[p] = ST(LD(P[p]) - (dt * k) * (((dvx / dx) + (dvy / dy)) + (dvz / dz)));                                          \
      }

#define ID(x) (x)
void pso_acoustic3d_v_f64(int nx, int ny, int nz, double* Vx, double* Vy, double* Vz, const double* P, double dt, double rho,
                          double dx, double dy, double dz) {
#pragma omp parallel for schedule(static)
  ACOUSTIC3D_V_BODY(ID, ID, double)
}
void pso_acoustic3d_p_f64(int nx, int ny, int nz, double* P, const double* Vx, const double* Vy, const double* Vz, double dt,
                          double k, double dx, double dy, double dz) {
#pragma omp parallel for schedule(static)
  ACOUSTIC3D_P_BODY(ID, ID, double)
}
void pso_acoustic3d_v_f32(int nx, int ny, int nz, float* Vx, float* Vy, float* Vz, const float* P, float dt, float rho, float dx,
                          float dy, float dz) {
#pragma omp parallel for schedule(static)
  ACOUSTIC3D_V_BODY(ID, ID, float)
}
void pso_acoustic3d_p_f32(int nx, int ny, int nz, float* P, const float* Vx, const float* Vy, const float* Vz, float dt, float k,
                          float dx, float dy, float dz) {
#pragma omp parallel for schedule(static)
  ACOUSTIC3D_P_BODY(ID, ID, float)
}
/* BFloat16 storage, Float32 arithmetic: loads widen, stores round to nearest even. */
void pso_acoustic3d_v_bf16(int nx, int ny, int nz, uint16_t* Vx, uint16_t* Vy, uint16_t* Vz, const uint16_t* P, float dt, float rho,
                           float dx, float dy, float dz) {
#pragma omp parallel for schedule(static)
  ACOUSTIC3D_V_BODY(bf16_ld, bf16_st, float)
}
void pso_acoustic3d_p_bf16(int nx, int ny, int nz, uint16_t* P, const uint16_t* Vx, const uint16_t* Vy, const uint16_t* Vz,
                           float dt, float k, float dx, float dy, float dz) {
#pragma omp parallel for schedule(static)
  ACOUSTIC3D_P_BODY(bf16_ld, bf16_st, float)
}

/* ------------------------------------------------------------ acoustic 2-D */
/* miniapps/acoustic2D.jl:11-20 (ranges 1:(n+1)). P: nx*ny, Vx: (nx+1)*ny, Vy: nx*(ny+1). */
void pso_acoustic2d_v_f64(int nx, int ny, double* Vx, double* Vy, const double* P, double dt, double rho, double dx, double dy) {
#pragma omp parallel for schedule(static)
  for (int iy = 1; iy <= ny + 1; ++iy)
    for (int ix = 1; ix <= nx + 1; ++ix) {
      if (1 <= ix && ix <= nx - 1 && 1 <= iy && iy <= ny - 2) {
        size_t v = IX2(ix + 1, iy + 1, nx + 1);
        Vx[v] = Vx[v] - ((dt / rho) * (P[IX2(ix + 1, iy + 1, nx)] - P[IX2(ix, iy + 1, nx)])) / dx;
      }
      if (1 <= ix && ix <= nx - 2 && 1 <= iy && iy <= ny - 1) {
        size_t v = IX2(ix + 1, iy + 1, nx);
        Vy[v] = Vy[v] - ((dt / rho) * (P[IX2(ix + 1, iy + 1, nx)] - P[IX2(ix + 1, iy, nx)])) / dy;
      }
    }
}
void pso_acoustic2d_p_f64(int nx, int ny, double* P, const double* Vx, const double* Vy, double dt, double k, double dx, double dy) {
#pragma omp parallel for schedule(static)
  for (int iy = 1; iy <= ny + 1; ++iy)
    for (int ix = 1; ix <= nx + 1; ++ix) {
      if (!(1 <= ix && ix <= nx && 1 <= iy && iy <= ny)) continue;
      double dvx = Vx[IX2(ix + 1, iy, nx + 1)] - Vx[IX2(ix, iy, nx + 1)];
      double dvy = Vy[IX2(ix, iy + 1, nx)] - Vy[IX2(ix, iy, nx)];
      size_t p = IX2(ix, iy, nx);
      P[p] = P[p] - (dt * k) * ((dvx / dx) + (dvy / dy));
    }
}

/* -------------------------------------------------------------- Stokes 2-D */
/* miniapps/Stokes2D.jl:11-55. Sizes (nx, ny = cells): Pt, dτPt, ∇V, τxx, τyy, Mus, Rog: nx*ny;
 * Vx: (nx+1)*ny; Vy: nx*(ny+1); τxy: (nx-1)*(ny-1); Rx, dVxdτ, dτVx: (nx-1)*(ny-2);
 * Ry, dVydτ, dτVy: (nx-2)*(ny-1). */
void pso_stokes2d_timesteps(int nx, int ny, double* dtVx, double* dtVy, double* dtPt, const double* Mus, double Vsc, double Ptsc,
                            double min_dxy2, int64_t max_nxy) {
#pragma omp parallel for schedule(static)
  for (int iy = 1; iy <= ny; ++iy)
    for (int ix = 1; ix <= nx; ++ix) {
      if (ix <= nx - 1 && iy <= ny - 2) {
        double av = (Mus[IX2(ix, iy + 1, nx)] + Mus[IX2(ix + 1, iy + 1, nx)]) * 0.5;
        dtVx[IX2(ix, iy, nx - 1)] = ((Vsc * min_dxy2) / av) / 4.1;
      }
      if (ix <= nx - 2 && iy <= ny - 1) {
        double av = (Mus[IX2(ix + 1, iy, nx)] + Mus[IX2(ix + 1, iy + 1, nx)]) * 0.5;
        dtVy[IX2(ix, iy, nx - 2)] = ((Vsc * min_dxy2) / av) / 4.1;
      }
      dtPt[IX2(ix, iy, nx)] = ((Ptsc * 4.1) * Mus[IX2(ix, iy, nx)]) / (double)max_nxy;
    }
}

void pso_stokes2d_p(int nx, int ny, double* divV, double* Pt, const double* Vx, const double* Vy, const double* dtPt, double dx,
                    double dy) {
#pragma omp parallel for schedule(static)
  for (int iy = 1; iy <= ny + 1; ++iy)
    for (int ix = 1; ix <= nx + 1; ++ix) {
      if (!(ix <= nx && iy <= ny)) continue;
      size_t c = IX2(ix, iy, nx);
      divV[c] = ((Vx[IX2(ix + 1, iy, nx + 1)] - Vx[IX2(ix, iy, nx + 1)]) / dx) + ((Vy[IX2(ix, iy + 1, nx)] - Vy[IX2(ix, iy, nx)]) / dy);
      Pt[c] = Pt[c] - dtPt[c] * divV[c];
    }
}

void pso_stokes2d_tau(int nx, int ny, const double* divV, double* txx, double* tyy, double* txy, const double* Vx, const double* Vy,
                      const double* Mus, double dx, double dy) {
#pragma omp parallel for schedule(static)
  for (int iy = 1; iy <= ny + 1; ++iy)
    for (int ix = 1; ix <= nx + 1; ++ix) {
      if (ix <= nx && iy <= ny) {
        size_t c = IX2(ix, iy, nx);
        double dvx = (Vx[IX2(ix + 1, iy, nx + 1)] - Vx[IX2(ix, iy, nx + 1)]) / dx;
        txx[c] = (2.0 * Mus[c]) * (dvx - ((1.0 / 3.0) * divV[c]));
        double dvy = (Vy[IX2(ix, iy + 1, nx)] - Vy[IX2(ix, iy, nx)]) / dy;
        tyy[c] = (2.0 * Mus[c]) * (dvy - ((1.0 / 3.0) * divV[c]));
      }
      if (ix <= nx - 1 && iy <= ny - 1) {
        double av = (((Mus[IX2(ix, iy, nx)] + Mus[IX2(ix + 1, iy, nx)]) + Mus[IX2(ix, iy + 1, nx)]) + Mus[IX2(ix + 1, iy + 1, nx)]) * 0.25;
        double dyiVx = (Vx[IX2(ix + 1, iy + 1, nx + 1)] - Vx[IX2(ix + 1, iy, nx + 1)]) / dy;
        double dxiVy = (Vy[IX2(ix + 1, iy + 1, nx)] - Vy[IX2(ix, iy + 1, nx)]) / dx;
        txy[IX2(ix, iy, nx - 1)] = (2.0 * av) * (0.5 * (dyiVx + dxiVy));
      }
    }
}

void pso_stokes2d_dv(int nx, int ny, double* Rx, double* Ry, double* dVxdt, double* dVydt, const double* Pt, const double* Rog,
                     const double* txx, const double* tyy, const double* txy, double dampX, double dampY, double dx, double dy) {
#pragma omp parallel for schedule(static)
  for (int iy = 1; iy <= ny + 1; ++iy)
    for (int ix = 1; ix <= nx + 1; ++ix) {
      if (ix <= nx - 1 && iy <= ny - 2) {
        size_t r = IX2(ix, iy, nx - 1);
        double a = (txx[IX2(ix + 1, iy + 1, nx)] - txx[IX2(ix, iy + 1, nx)]) / dx;
        double b = (txy[IX2(ix, iy + 1, nx - 1)] - txy[IX2(ix, iy, nx - 1)]) / dy;
        double c = (Pt[IX2(ix + 1, iy + 1, nx)] - Pt[IX2(ix, iy + 1, nx)]) / dx;
        Rx[r] = (a + b) - c;
      }
      if (ix <= nx - 2 && iy <= ny - 1) {
        size_t r = IX2(ix, iy, nx - 2);
        double a = (tyy[IX2(ix + 1, iy + 1, nx)] - tyy[IX2(ix + 1, iy, nx)]) / dy;
        double b = (txy[IX2(ix + 1, iy, nx - 1)] - txy[IX2(ix, iy, nx - 1)]) / dx;
        double c = (Pt[IX2(ix + 1, iy + 1, nx)] - Pt[IX2(ix + 1, iy, nx)]) / dy;
        double av = (Rog[IX2(ix + 1, iy, nx)] + Rog[IX2(ix + 1, iy + 1, nx)]) * 0.5;
        Ry[r] = ((a + b) - c) + av;
      }
      if (ix <= nx - 1 && iy <= ny - 2) {
        size_t r = IX2(ix, iy, nx - 1);
        dVxdt[r] = dampX * dVxdt[r] + Rx[r];
      }
      if (ix <= nx - 2 && iy <= ny - 1) {
        size_t r = IX2(ix, iy, nx - 2);
        dVydt[r] = dampY * dVydt[r] + Ry[r];
      }
    }
}

void pso_stokes2d_v(int nx, int ny, double* Vx, double* Vy, const double* dVxdt, const double* dVydt, const double* dtVx,
                    const double* dtVy) {
#pragma omp parallel for schedule(static)
  for (int iy = 1; iy <= ny + 1; ++iy)
    for (int ix = 1; ix <= nx + 1; ++ix) {
      if (1 <= ix && ix <= nx - 1 && 1 <= iy && iy <= ny - 2) {
        size_t v = IX2(ix + 1, iy + 1, nx + 1);
        Vx[v] = Vx[v] + dtVx[IX2(ix, iy, nx - 1)] * dVxdt[IX2(ix, iy, nx - 1)];
      }
      if (1 <= ix && ix <= nx - 2 && 1 <= iy && iy <= ny - 1) {
        size_t v = IX2(ix + 1, iy + 1, nx);
        Vy[v] = Vy[v] + dtVy[IX2(ix, iy, nx - 2)] * dVydt[IX2(ix, iy, nx - 2)];
      }
    }
}

/* bc_x!(A): A[1,iy] = A[2,iy]; A[end,iy] = A[end-1,iy]  (miniapps/Stokes2D.jl:44-48) */
void pso_bc_x(int n1, int n2, double* A) {
  for (int iy = 1; iy <= n2; ++iy) {
    A[IX2(1, iy, n1)] = A[IX2(2, iy, n1)];
    A[IX2(n1, iy, n1)] = A[IX2(n1 - 1, iy, n1)];
  }
}
/* bc_y!(A): A[ix,1] = A[ix,2]; A[ix,end] = A[ix,end-1]  (miniapps/Stokes2D.jl:50-54) */
void pso_bc_y(int n1, int n2, double* A) {
  for (int ix = 1; ix <= n1; ++ix) {
    A[IX2(ix, 1, n1)] = A[IX2(ix, 2, n1)];
    A[IX2(ix, n2, n1)] = A[IX2(ix, n2 - 1, n1)];
  }
}

/* mean(abs.(A)) over n values, summed in index order (miniapps/Stokes2D.jl:119). */
double pso_mean_abs(int64_t n, const double* A) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += fabs(A[i]);
  return n ? s / (double)n : 0.0;
}
