// convolve.hip -- auralization: convolution of impulse-response tracks with
// a sound signal (RecorderTrack::Process, src/Recorder.cpp:145-243).
//
// The reference transforms the zero-padded track a and signal b (both of
// length M + N + 1) with FFTW, multiplies the spectra, scales by 1 / (M + N + 1)
// and transforms back; with a second track (keyframe interpolation) it adds
// the convolution of that track with the signal faded in to the convolution
// of the first with the signal faded out.  Here the transforms are batched
// hipFFT real transforms of a power-of-two length L >= M + N + 1 (more zero
// padding leaves the linear convolution unchanged); both products are summed
// in the frequency domain, so one inverse transform serves both terms.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "ear_internal.hpp"

namespace ear {
namespace {

// C[k] = (A0[k] B0[k] + A1[k] B1[k]) * scale over nc bins; terms = 1 or 2
__global__ void k_spectra_mul(const float2* __restrict__ X, uint32_t nc, uint32_t terms, float scale,
                              float2* __restrict__ C) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nc; k += gridDim.x * blockDim.x) {
        float re = 0.0f, im = 0.0f;
        for (uint32_t t = 0; t < terms; ++t) {
            const float2 a = X[(size_t)t * nc + k];                // track t
            const float2 b = X[((size_t)terms + t) * nc + k];      // signal t
            re += a.x * b.x - a.y * b.y;
            im += a.x * b.y + a.y * b.x;
        }
        C[k] = make_float2(re * scale, im * scale);
    }
}

struct Plan {
    hipfftHandle h = 0;
    ~Plan() {
        if (h) hipfftDestroy(h);
    }
};

ear_status fft_status(Ctx* c, hipfftResult r, const char* what) {
    if (r == HIPFFT_SUCCESS) return EAR_OK;
    c->last_error = std::string(what) + " failed (hipFFT error " + std::to_string((int)r) + ")";
    return r == HIPFFT_ALLOC_FAILED ? EAR_ERR_OUT_OF_MEMORY : EAR_ERR_HIP;
}

}  // namespace

ear_status run_convolve(Ctx* c, const float* const* tracks, const uint32_t* n_track, const float* const* signals,
                        uint32_t terms, uint32_t n_sig, float* out, uint32_t n_out) {
    uint64_t need = n_out;
    uint32_t L = 1;
    while (L < need) L <<= 1;
    const uint32_t nc = L / 2 + 1;
    hipStream_t st = c->stream;
    DevBuf real(c), spec(c), prod(c);
    // inputs: tracks then signals, each zero-padded to L reals
    TRY(real.alloc((size_t)2 * terms * L * sizeof(float)));
    TRY(spec.alloc((size_t)2 * terms * nc * sizeof(float2)));
    TRY(prod.alloc((size_t)nc * sizeof(float2)));
    EAR_HIP_TRY(c, hipMemsetAsync(real.p, 0, real.n, st));
    for (uint32_t t = 0; t < terms; ++t) {
        EAR_HIP_TRY(c, hipMemcpyAsync(real.as<float>() + (size_t)t * L, tracks[t], (size_t)n_track[t] * 4,
                                      hipMemcpyHostToDevice, st));
        EAR_HIP_TRY(c, hipMemcpyAsync(real.as<float>() + (size_t)(terms + t) * L, signals[t], (size_t)n_sig * 4,
                                      hipMemcpyHostToDevice, st));
    }
    Plan fwd, inv;
    int n = (int)L;
    int inembed = (int)L, onembed = (int)nc;
    TRY(fft_status(c, hipfftPlanMany(&fwd.h, 1, &n, &inembed, 1, (int)L, &onembed, 1, (int)nc, HIPFFT_R2C,
                                     (int)(2 * terms)),
                   "hipfftPlanMany (forward)"));
    TRY(fft_status(c, hipfftPlan1d(&inv.h, n, HIPFFT_C2R, 1), "hipfftPlan1d (inverse)"));
    TRY(fft_status(c, hipfftSetStream(fwd.h, st), "hipfftSetStream"));
    TRY(fft_status(c, hipfftSetStream(inv.h, st), "hipfftSetStream"));
    TRY(fft_status(c, hipfftExecR2C(fwd.h, real.as<hipfftReal>(), spec.as<hipfftComplex>()), "forward FFT"));
    k_spectra_mul<<<std::min<uint32_t>((nc + 255) / 256, 4096), 256, 0, st>>>(spec.as<float2>(), nc, terms,
                                                                            1.0f / (float)L, prod.as<float2>());
    EAR_HIP_TRY(c, hipGetLastError());
    TRY(fft_status(c, hipfftExecC2R(inv.h, prod.as<hipfftComplex>(), real.as<hipfftReal>()), "inverse FFT"));
    EAR_HIP_TRY(c, hipMemcpyAsync(out, real.p, (size_t)n_out * 4, hipMemcpyDeviceToHost, st));
    EAR_HIP_TRY(c, hipStreamSynchronize(st));
    return EAR_OK;
}

}  // namespace ear
