// Host dense kernels of the Krylov exponential's small problems.
#include <complex>
#include <cstddef>

namespace pimi {

// C = A B for q x q column-major complex matrices held as real and imaginary
// planes (A, C) and interleaved pairs (B): the innermost loop runs down a
// column with unit stride, so it vectorizes; built for AVX2 + FMA and for the
// baseline ISA, the loader picks one.  With std::complex arithmetic in the
// innermost loop a q = 41 exponential took 0.9 ms.
__attribute__((target_clones("avx2,fma", "default"))) void mm_planes(int q, const double *__restrict__ ar, const double *__restrict__ ai,
                                                                      const std::complex<double> *__restrict__ B,
                                                                      double *__restrict__ cr, double *__restrict__ ci) {
    for (int j = 0; j < q; ++j) {
        double *__restrict__ crj = cr + (size_t)j * q;
        double *__restrict__ cij = ci + (size_t)j * q;
        for (int i = 0; i < q; ++i) {
            crj[i] = 0;
            cij[i] = 0;
        }
        for (int k = 0; k < q; ++k) {
            const double br = B[(size_t)k + (size_t)j * q].real(), bi = B[(size_t)k + (size_t)j * q].imag();
            if (br == 0.0 && bi == 0.0) continue;
            const double *__restrict__ ark = ar + (size_t)k * q;
            const double *__restrict__ aik = ai + (size_t)k * q;
            for (int i = 0; i < q; ++i) {
                crj[i] += ark[i] * br - aik[i] * bi;
                cij[i] += ark[i] * bi + aik[i] * br;
            }
        }
    }
}

}  // namespace pimi
