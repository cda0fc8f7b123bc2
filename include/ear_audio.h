/* ear_audio.h -- C ABI of the audio side of auralization, exported by the
 * host library (libear_host.so): WAV files and the three-band equalizer that
 * splits a sound source before its bands are convolved with the impulse
 * responses (ear_convolve in ear.h).  Host code only; no device call.
 *
 * Return codes: 0 on success, -1 when the file cannot be opened, read or
 * written or is not a supported WAV file, -2 on a null pointer, -3 when the
 * output buffer is too small (n then holds the frames needed). */
#ifndef EAR_AUDIO_H
#define EAR_AUDIO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* WaveFile::Load + WaveFile::ToFloat (lib/wave/WaveFile.h:49-75,
 * lib/wave/WaveFile.cpp:56-189): integer PCM of 8, 16 or 24 bits, the
 * channels of each frame averaged, samples divided by 2^(bits - 1) (8-bit
 * samples are unsigned around 128).  Writes up to cap frames to out (out may
 * be NULL with cap 0 to query the length), the frame count to *n and, when
 * not NULL, the file's channel count and bit depth. */
int ear_wave_read(const char* path, float* out, uint64_t cap, uint64_t* n, int* channels, int* bits);

/* WaveFile::FromFloat + WaveFile::Save (lib/wave/WaveFile.cpp:116-256):
 * 16-bit PCM at 44.1 kHz, sample = x / max * 32768 truncated toward zero.
 * right == NULL writes one channel: max = 1, or with norm != 0 the largest
 * |x| / 0.8, or norm_max / 0.95 when norm_max >= 0.  With right, two
 * interleaved channels, the shorter padded with zeros; with norm != 0 max is
 * the largest signed sample of both / 0.8 (norm_max unused).  Samples beyond
 * the 16-bit range saturate. */
int ear_wave_write(const char* path, const float* left, uint64_t n_left, const float* right, uint64_t n_right,
                   int norm, float norm_max);

/* Equalizer::Split (lib/equalizer/Equalizer.h:26-49,
 * lib/equalizer/Equalizer.cpp:81-95): fourth-order Linkwitz-Riley crossovers
 * at fc1 = (f1 + f2) / 2 and fc2 = (f2 + f3) / 2 (Hz, fs = 44.1 kHz):
 * low = LP(fc1) x, mid = HP(fc1) LP(fc2) x, high = HP(fc2) x, each filter a
 * single-precision direct-form-I recursion from rest. */
int ear_equalizer_split(const float* x, uint64_t n, float* low, float* mid, float* high, float f1, float f2,
                        float f3);

#ifdef __cplusplus
}
#endif

#endif /* EAR_AUDIO_H */
