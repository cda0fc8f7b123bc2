// audio.h -- audio side of `EAR render`: WAV files and the three-band
// split of single-file sound sources.
#pragma once

#include <string>
#include <vector>

namespace earhost {

// WaveFile (lib/wave/WaveFile.cpp:56-256): PCM RIFF files.
//
// read: 8-, 16- or 24-bit integer PCM; the channels of a frame are averaged
// and samples scaled by 2^(bits-1) (8-bit samples are unsigned around 128),
// as WaveFile::ToFloat.  Returns false on a missing or unsupported file.
bool wave_read(const std::string& path, std::vector<float>* samples, int* channels = nullptr,
               int* bits = nullptr);

// write: 16-bit PCM at 44.1 kHz, sample = (short)(x / max * 32768) with
// max = 1 unless normalizing (WaveFile::FromFloat).  Mono: with norm the
// largest |x| maps to 0.8, or norm_max / 0.95 when norm_max >= 0.  Stereo:
// interleaved, the shorter channel padded with zeros; with norm the largest
// signed sample maps to 0.8.
bool wave_write_mono(const std::string& path, const float* x, size_t n, bool norm = false, float norm_max = -1.0f);
bool wave_write_stereo(const std::string& path, const float* left, size_t n_left, const float* right, size_t n_right,
                       bool norm = false);

// Equalizer::Split (lib/equalizer/Equalizer.cpp:81-95): fourth-order
// Linkwitz-Riley crossovers at fc1 = (f1 + f2) / 2 and fc2 = (f2 + f3) / 2
// (Hz): low = LP(fc1) x, mid = HP(fc1) LP(fc2) x, high = HP(fc2) x.
void equalizer_split(const float* x, size_t n, float* low, float* mid, float* high, float f1, float f2, float f3);

// One fourth-order Linkwitz-Riley section (a squared second-order
// Butterworth) by the bilinear transform, run as a direct-form-I recursion in
// single precision.
class LinkwitzRiley4 {
   public:
    LinkwitzRiley4(float fc, bool high_pass);
    float process(float in);
    float a[5], b[5];  // numerator and denominator (b[0] = 1)

   private:
    float xm[4] = {0, 0, 0, 0}, ym[4] = {0, 0, 0, 0};
};

}  // namespace earhost
