// Small dense solvers used inside the per-cell flash (one GPU lane per cell).
//
// The reference solves its Newton systems with LAPACK: dsysv with UPLO='U'
// (src/DenseMat.cpp:234-253; Bunch-Kaufman LDL^T, the published reference-
// LAPACK algorithm of dsytf2 + dsytrs) and dgesv (src/DenseMat.cpp:221-232;
// partial-pivoting LU, dgetf2 + dgetrs).  Both are restated here for a
// single lane with compile-time sizes so that the factors live in registers.
//
// Symmetric storage: the reference hands LAPACK a C row-major array, which
// LAPACK reads column-major; UPLO='U' then touches only entries (r, c) with
// r <= c in LAPACK's view, i.e. the C-lower triangle.  `SymPacked<N>` stores
// exactly those N(N+1)/2 values: U(r, c) (r <= c) at c*(c+1)/2 + r, which is
// also the row-major packed lower triangle of the C array (row c, column r).
#pragma once

#include <hip/hip_runtime.h>

namespace ocpg {

template <int N>
struct SymPacked {
    static constexpr int kSize = N * (N + 1) / 2;
    double a[kSize];
    __device__ __forceinline__ static constexpr int idx(int r, int c) { return c * (c + 1) / 2 + r; }
    __device__ __forceinline__ double& U(int r, int c) { return a[idx(r, c)]; }
};

// dsytf2 (UPLO='U') followed by dsytrs for one right-hand side.  Returns the
// LAPACK info value; on a singular factor (info > 0) b is left unsolved, as
// dsysv leaves it.
template <int N>
__device__ __forceinline__ int symIndefSolve(SymPacked<N>& A, double* b)
{
    const double alpha = (1.0 + sqrt(17.0)) / 8.0;
    int ipiv[N];
    int info = 0;
    bool consumed = false;  // column already used by a 2x2 pivot
#pragma unroll
    for (int k = N - 1; k >= 0; --k) {
        if (consumed) {
            consumed = false;
            continue;
        }
        int kstep = 1;
        const double absakk = fabs(A.U(k, k));
        int imax = 0;
        double colmax = 0.0;
        if (k > 0) {
            colmax = fabs(A.U(0, k));
            for (int i = 1; i < k; ++i) {
                const double v = fabs(A.U(i, k));
                if (v > colmax) {
                    colmax = v;
                    imax = i;
                }
            }
        }
        int kp = k;
        if (fmax(absakk, colmax) == 0.0) {
            if (info == 0) info = k + 1;
        } else {
            if (!(absakk >= alpha * colmax)) {
                // largest off-diagonal of row/column imax
                double rowmax = fabs(A.U(imax, imax + 1));
                for (int j = imax + 2; j <= k; ++j) rowmax = fmax(rowmax, fabs(A.U(imax, j)));
                if (imax > 0) {
                    double cm = fabs(A.U(0, imax));
                    for (int i = 1; i < imax; ++i) cm = fmax(cm, fabs(A.U(i, imax)));
                    rowmax = fmax(rowmax, cm);
                }
                if (absakk >= alpha * colmax * (colmax / rowmax)) {
                    kp = k;
                } else if (fabs(A.U(imax, imax)) >= alpha * rowmax) {
                    kp = imax;
                } else {
                    kp = imax;
                    kstep = 2;
                }
            }
            const int kk = k - kstep + 1;
            if (kp != kk) {
                for (int i = 0; i < kp; ++i) {
                    const double t = A.U(i, kk);
                    A.U(i, kk) = A.U(i, kp);
                    A.U(i, kp) = t;
                }
                for (int j = kp + 1; j < kk; ++j) {
                    const double t = A.U(j, kk);
                    A.U(j, kk) = A.U(kp, j);
                    A.U(kp, j) = t;
                }
                {
                    const double t = A.U(kk, kk);
                    A.U(kk, kk) = A.U(kp, kp);
                    A.U(kp, kp) = t;
                }
                if (kstep == 2) {
                    const double t = A.U(k - 1, k);
                    A.U(k - 1, k) = A.U(kp, k);
                    A.U(kp, k) = t;
                }
            }
            if (kstep == 1) {
                const double r1 = 1.0 / A.U(k, k);
                const double malpha = -r1;
#pragma unroll
                for (int j = 0; j < k; ++j) {
                    if (A.U(j, k) != 0.0) {
                        const double temp = malpha * A.U(j, k);
#pragma unroll
                        for (int i = 0; i <= j; ++i) A.U(i, j) = A.U(i, j) + A.U(i, k) * temp;
                    }
                }
#pragma unroll
                for (int i = 0; i < k; ++i) A.U(i, k) = r1 * A.U(i, k);
            } else if (k > 1) {
                double d12 = A.U(k - 1, k);
                const double d22 = A.U(k - 1, k - 1) / d12;
                const double d11 = A.U(k, k) / d12;
                const double t = 1.0 / (d11 * d22 - 1.0);
                d12 = t / d12;
#pragma unroll
                for (int j = k - 2; j >= 0; --j) {
                    const double wkm1 = d12 * (d11 * A.U(j, k - 1) - A.U(j, k));
                    const double wk = d12 * (d22 * A.U(j, k) - A.U(j, k - 1));
#pragma unroll
                    for (int i = j; i >= 0; --i) A.U(i, j) = A.U(i, j) - A.U(i, k) * wk - A.U(i, k - 1) * wkm1;
                    A.U(j, k) = wk;
                    A.U(j, k - 1) = wkm1;
                }
            }
        }
        if (kstep == 1) {
            ipiv[k] = kp + 1;
        } else {
            ipiv[k] = -(kp + 1);
            if (k > 0) ipiv[k - 1] = -(kp + 1);
            consumed = true;
        }
    }
    if (info != 0) return info;
    // dsytrs: solve U D X = b
    consumed = false;
#pragma unroll
    for (int k = N - 1; k >= 0; --k) {
        if (consumed) {
            consumed = false;
            continue;
        }
        if (ipiv[k] > 0) {
            const int kp = ipiv[k] - 1;
            if (kp != k) {
                const double t = b[k];
                b[k] = b[kp];
                b[kp] = t;
            }
            const double bk = -b[k];
            if (bk != 0.0) {
#pragma unroll
                for (int i = 0; i < k; ++i) b[i] = b[i] + A.U(i, k) * bk;
            }
            b[k] = (1.0 / A.U(k, k)) * b[k];
        } else if (k > 0) {
            const int kp = -ipiv[k] - 1;
            if (kp != k - 1) {
                const double t = b[k - 1];
                b[k - 1] = b[kp];
                b[kp] = t;
            }
            double t = -b[k];
            if (t != 0.0) {
#pragma unroll
                for (int i = 0; i < k - 1; ++i) b[i] = b[i] + A.U(i, k) * t;
            }
            t = -b[k - 1];
            if (t != 0.0) {
#pragma unroll
                for (int i = 0; i < k - 1; ++i) b[i] = b[i] + A.U(i, k - 1) * t;
            }
            const double akm1k = A.U(k - 1, k);
            const double akm1 = A.U(k - 1, k - 1) / akm1k;
            const double ak = A.U(k, k) / akm1k;
            const double denom = akm1 * ak - 1.0;
            const double bkm1 = b[k - 1] / akm1k;
            const double bkk = b[k] / akm1k;
            b[k - 1] = (ak * bkm1 - bkk) / denom;
            b[k] = (akm1 * bkk - bkm1) / denom;
            consumed = true;
        }
    }
    // solve U^T X = b
    consumed = false;
#pragma unroll
    for (int k = 0; k < N; ++k) {
        if (consumed) {
            consumed = false;
            continue;
        }
        if (ipiv[k] > 0) {
            double s = 0.0;
#pragma unroll
            for (int i = 0; i < k; ++i) s += A.U(i, k) * b[i];
            b[k] = b[k] - s;
            const int kp = ipiv[k] - 1;
            if (kp != k) {
                const double t = b[k];
                b[k] = b[kp];
                b[kp] = t;
            }
        } else if (k + 1 < N) {
            double s = 0.0;
#pragma unroll
            for (int i = 0; i < k; ++i) s += A.U(i, k) * b[i];
            b[k] = b[k] - s;
            s = 0.0;
#pragma unroll
            for (int i = 0; i < k; ++i) s += A.U(i, k + 1) * b[i];
            b[k + 1] = b[k + 1] - s;
            const int kp = -ipiv[k] - 1;
            if (kp != k) {
                const double t = b[k];
                b[k] = b[kp];
                b[kp] = t;
            }
            consumed = true;
        }
    }
    return 0;
}

// dgetf2 + dgetrs on an N x N column-major matrix with R right-hand sides
// (column-major N x R).  Returns info; on a singular factor b is unsolved.
template <int N, int R>
__device__ __forceinline__ int luSolve(double* a, double* b)
{
#define LA(r, c) a[(r) + (c)*N]
    int ipiv[N];
    int info = 0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        int p = j;
        double pm = fabs(LA(j, j));
#pragma unroll
        for (int i = j + 1; i < N; ++i) {
            const double v = fabs(LA(i, j));
            if (v > pm) {
                pm = v;
                p = i;
            }
        }
        ipiv[j] = p;
        if (LA(p, j) != 0.0) {
            if (p != j) {
#pragma unroll
                for (int c = 0; c < N; ++c) {
                    const double t = LA(j, c);
                    LA(j, c) = LA(p, c);
                    LA(p, c) = t;
                }
            }
            const double r = 1.0 / LA(j, j);
#pragma unroll
            for (int i = j + 1; i < N; ++i) LA(i, j) = r * LA(i, j);
        } else if (info == 0) {
            info = j + 1;
        }
#pragma unroll
        for (int c = j + 1; c < N; ++c) {
            const double t = LA(j, c);
            if (t != 0.0) {
#pragma unroll
                for (int i = j + 1; i < N; ++i) LA(i, c) = LA(i, c) - LA(i, j) * t;
            }
        }
    }
    if (info != 0) return info;
#pragma unroll
    for (int c = 0; c < R; ++c) {
        double* bc = b + c * N;
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const int p = ipiv[j];
            if (p != j) {
                const double t = bc[j];
                bc[j] = bc[p];
                bc[p] = t;
            }
        }
#pragma unroll
        for (int j = 0; j < N; ++j) {
            const double t = bc[j];
            if (t != 0.0) {
#pragma unroll
                for (int i = j + 1; i < N; ++i) bc[i] = bc[i] - t * LA(i, j);
            }
        }
#pragma unroll
        for (int j = N - 1; j >= 0; --j) {
            if (bc[j] != 0.0) {
                bc[j] = bc[j] / LA(j, j);
                const double t = bc[j];
#pragma unroll
                for (int i = 0; i < j; ++i) bc[i] = bc[i] - t * LA(i, j);
            }
        }
    }
#undef LA
    return 0;
}

// Smallest eigenvalue of a symmetric N x N matrix given in FP32 (what the
// reference's ssyevd returns first, src/DenseMat.cpp:14-26): cyclic Jacobi
// in FP64 on the FP32 entries, rounded to FP32.
template <int N>
__device__ __forceinline__ float minEigenSym(const float* af)
{
    double a[N * N];
#pragma unroll
    for (int i = 0; i < N * N; ++i) a[i] = af[i];
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0;
#pragma unroll
        for (int p = 0; p < N; ++p)
#pragma unroll
            for (int q = p + 1; q < N; ++q) off += a[p * N + q] * a[p * N + q];
        if (off < 1e-300) break;
#pragma unroll
        for (int p = 0; p < N; ++p) {
#pragma unroll
            for (int q = p + 1; q < N; ++q) {
                const double apq = a[p * N + q];
                if (apq != 0.0) {
                    const double app = a[p * N + p];
                    const double aqq = a[q * N + q];
                    const double theta = (aqq - app) / (2.0 * apq);
                    const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                    const double c = 1.0 / sqrt(t * t + 1.0);
                    const double s = t * c;
#pragma unroll
                    for (int k = 0; k < N; ++k) {
                        const double akp = a[k * N + p];
                        const double akq = a[k * N + q];
                        a[k * N + p] = c * akp - s * akq;
                        a[k * N + q] = s * akp + c * akq;
                    }
#pragma unroll
                    for (int k = 0; k < N; ++k) {
                        const double apk = a[p * N + k];
                        const double aqk = a[q * N + k];
                        a[p * N + k] = c * apk - s * aqk;
                        a[q * N + k] = s * apk + c * aqk;
                    }
                }
            }
        }
    }
    double m = a[0];
#pragma unroll
    for (int i = 1; i < N; ++i) m = fmin(m, a[i * N + i]);
    return static_cast<float>(m);
}

}  // namespace ocpg
