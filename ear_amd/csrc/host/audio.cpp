// audio.cpp -- WAV files and the three-band equalizer of `EAR render`.
#include "audio.h"

#include "ear_audio.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

namespace earhost {

namespace {

uint32_t le32(const uint8_t* p) { return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24; }
uint16_t le16(const uint8_t* p) { return (uint16_t)(p[0] | p[1] << 8); }

void put32(std::vector<uint8_t>& o, uint32_t v) {
    for (int i = 0; i < 4; ++i) o.push_back((uint8_t)(v >> (8 * i)));
}
void put16(std::vector<uint8_t>& o, uint16_t v) {
    o.push_back((uint8_t)v);
    o.push_back((uint8_t)(v >> 8));
}
void put_id(std::vector<uint8_t>& o, const char* id) { o.insert(o.end(), id, id + 4); }

// A float sample scaled to 16 bits: truncation toward zero as the
// reference's (short) cast; values outside the 16-bit range, undefined there,
// saturate here and NaN becomes 0.
int16_t to_pcm16(float v) {
    if (!(v == v)) return 0;
    if (v >= 32767.0f) return 32767;
    if (v <= -32768.0f) return -32768;
    return (int16_t)(int32_t)v;
}

// RIFF header of a 16-bit PCM file at 44.1 kHz (WaveFile::Save and the
// header fields FromFloat sets).
bool write_pcm16(const std::string& path, const std::vector<int16_t>& s, uint16_t channels) {
    const uint32_t data_bytes = (uint32_t)(s.size() * 2);
    std::vector<uint8_t> o;
    o.reserve(44 + data_bytes);
    put_id(o, "RIFF");
    put32(o, data_bytes + 36);
    put_id(o, "WAVE");
    put_id(o, "fmt ");
    put32(o, 16);
    put16(o, 1);  // integer PCM
    put16(o, channels);
    put32(o, 44100);
    put32(o, 44100u * 2u * channels);
    put16(o, (uint16_t)(2 * channels));
    put16(o, 16);
    put_id(o, "data");
    put32(o, data_bytes);
    for (int16_t v : s) put16(o, (uint16_t)v);
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(o.data(), 1, o.size(), f) == o.size();
    return std::fclose(f) == 0 && ok;
}

}  // namespace

bool wave_read(const std::string& path, std::vector<float>* samples, int* channels, int* bits) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::vector<uint8_t> b;
    uint8_t buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) b.insert(b.end(), buf, buf + got);
    std::fclose(f);
    if (b.size() < 12 || std::memcmp(b.data(), "RIFF", 4) != 0 || std::memcmp(b.data() + 8, "WAVE", 4) != 0)
        return false;
    // Chunks in file order; every "data" chunk is appended, as the
    // reference's loader does (lib/wave/WaveFile.cpp:84-101).  Chunk sizes are
    // honoured for every chunk (odd sizes padded to a word, as RIFF requires),
    // where the reference reads a fixed 16-byte "fmt " body.
    uint16_t tag = 0, ch = 0, bps = 0;
    bool have_fmt = false;
    std::vector<uint8_t> pcm;
    size_t p = 12;
    while (p + 8 <= b.size()) {
        const uint8_t* id = b.data() + p;
        const uint32_t sz = le32(b.data() + p + 4);
        p += 8;
        const size_t avail = std::min<size_t>(sz, b.size() - p);
        if (std::memcmp(id, "fmt ", 4) == 0 && avail >= 16) {
            tag = le16(b.data() + p);
            ch = le16(b.data() + p + 2);
            bps = le16(b.data() + p + 14);
            have_fmt = true;
        } else if (std::memcmp(id, "data", 4) == 0) {
            pcm.insert(pcm.end(), b.data() + p, b.data() + p + avail);
        }
        p += (size_t)sz + (sz & 1u);
    }
    if (!have_fmt || tag != 1 || ch == 0 || (bps != 8 && bps != 16 && bps != 24)) return false;
    const size_t bytes = bps / 8;
    const size_t frames = pcm.size() / bytes / ch;
    if (frames == 0) return false;
    // WaveFile::ToFloat (lib/wave/WaveFile.cpp:160-189): each channel sample
    // divided by 2^(bits-1) and summed, the sum divided by the channel count
    const float full = (float)(1u << (bps - 1));
    samples->assign(frames, 0.0f);
    const uint8_t* q = pcm.data();
    for (size_t i = 0; i < frames; ++i) {
        float acc = 0.0f;
        for (uint16_t c = 0; c < ch; ++c, q += bytes) {
            float v;
            if (bytes == 1)
                v = (float)q[0] - 128.0f;
            else if (bytes == 2)
                v = (float)(int16_t)le16(q);
            else
                v = (float)((int32_t)((uint32_t)q[0] << 8 | (uint32_t)q[1] << 16 | (uint32_t)q[2] << 24) >> 8);
            acc += v / full;
        }
        (*samples)[i] = acc / (float)ch;
    }
    if (channels) *channels = ch;
    if (bits) *bits = bps;
    return true;
}

bool wave_write_mono(const std::string& path, const float* x, size_t n, bool norm, float norm_max) {
    // WaveFile::FromFloat (lib/wave/WaveFile.cpp:191-221)
    float mx = 1.0f;
    if (norm) {
        if (norm_max < 0.0f) {
            mx = -1e9f;
            for (size_t i = 0; i < n; ++i) mx = std::max(mx, std::fabs(x[i]));
            mx /= 0.8f;
        } else {
            mx = norm_max / 0.95f;
        }
    }
    std::vector<int16_t> s(n);
    for (size_t i = 0; i < n; ++i) s[i] = to_pcm16(x[i] / mx * 32768.0f);
    return write_pcm16(path, s, 1);
}

bool wave_write_stereo(const std::string& path, const float* left, size_t n_left, const float* right, size_t n_right,
                       bool norm) {
    // WaveFile::FromFloat, two channels (lib/wave/WaveFile.cpp:223-256): the
    // normalisation takes the largest signed sample of either channel
    float mx = 1.0f;
    if (norm) {
        mx = -1e9f;
        for (size_t i = 0; i < n_left; ++i) mx = std::max(mx, left[i]);
        for (size_t i = 0; i < n_right; ++i) mx = std::max(mx, right[i]);
        mx /= 0.8f;
    }
    const size_t n = std::max(n_left, n_right);
    std::vector<int16_t> s(2 * n);
    for (size_t i = 0; i < n; ++i) {
        s[2 * i] = i < n_left ? to_pcm16(left[i] / mx * 32768.0f) : 0;
        s[2 * i + 1] = i < n_right ? to_pcm16(right[i] / mx * 32768.0f) : 0;
    }
    return write_pcm16(path, s, 2);
}

// Equalizer::Pass (lib/equalizer/Equalizer.cpp:27-79).  The analog prototype
// is a squared second-order Butterworth, (w^2 / (s^2 + sqrt(2) w s + w^2))^2
// for the low-pass and (s^2 / (...))^2 for the high-pass, with w = 2 pi fs;
// the bilinear map s = k (1 - z^-1) / (1 + z^-1), k = w / tan(pi fc / fs),
// puts the -6 dB point at fc.  Multiplying out (k^2 (1-z)^2 + sqrt2 w k
// (1-z)(1+z) + w^2 (1+z)^2)^2 gives the denominator terms below (published
// with this filter as musicdsp.org archive #266).  pi is the reference's
// 22/7, which moves the crossovers up by 0.04 %.  The coefficients are formed
// in single precision with the reference's order of operations: the filter's
// double poles make the output sensitive to the last bit of a coefficient
// (a one-ulp change moves a band by ~1e-3 of its peak), so only the same
// float sequence reproduces the reference's band signals.
LinkwitzRiley4::LinkwitzRiley4(float fc, bool high_pass) {
    const float fs = 44100.0f, pi = 3.14285714285714f;
    const float w = 2.0f * pi * fs;
    const float w2 = w * w, w3 = w2 * w, w4 = w2 * w2;
    const float k = w / std::tan(pi * fc / fs);
    const float k2 = k * k, k3 = k2 * k, k4 = k2 * k2;
    const float r2 = std::sqrt(2.0f);
    const float c1 = r2 * w3 * k, c3 = r2 * w * k3;  // the odd cross terms
    const float d0 = 4.0f * w2 * k2 + 2.0f * c1 + k4 + 2.0f * c3 + w4;
    b[0] = 1.0f;
    b[1] = (4.0f * (w4 + c1 - k4 - c3)) / d0;
    b[2] = (6.0f * w4 - 8.0f * w2 * k2 + 6.0f * k4) / d0;
    b[3] = (4.0f * (w4 - c1 + c3 - k4)) / d0;
    b[4] = (k4 - 2.0f * c1 + w4 - 2.0f * c3 + 4.0f * w2 * k2) / d0;
    // numerator g (1 + z)^4 (low-pass, g = w^4) or g (1 - z)^4 (high-pass, g = k^4)
    const float g = high_pass ? k4 : w4;
    a[0] = g / d0;
    a[1] = (high_pass ? -4.0f : 4.0f) * g / d0;
    a[2] = 6.0f * g / d0;
    a[3] = a[1];
    a[4] = a[0];
}

float LinkwitzRiley4::process(float in) {
    // direct form I, the terms summed in the reference's order
    const float out = a[0] * in + a[1] * xm[0] + a[2] * xm[1] + a[3] * xm[2] + a[4] * xm[3] - b[1] * ym[0] -
                      b[2] * ym[1] - b[3] * ym[2] - b[4] * ym[3];
    xm[3] = xm[2];
    xm[2] = xm[1];
    xm[1] = xm[0];
    xm[0] = in;
    ym[3] = ym[2];
    ym[2] = ym[1];
    ym[1] = ym[0];
    ym[0] = out;
    return out;
}

void equalizer_split(const float* x, size_t n, float* low, float* mid, float* high, float f1, float f2, float f3) {
    const float fc1 = (f1 + f2) / 2.0f, fc2 = (f2 + f3) / 2.0f;
    LinkwitzRiley4 hp1(fc1, true), hp2(fc2, true), lp1(fc1, false), lp2(fc2, false);
    for (size_t i = 0; i < n; ++i) {
        high[i] = hp2.process(x[i]);
        low[i] = lp1.process(x[i]);
        mid[i] = hp1.process(lp2.process(x[i]));
    }
}

}  // namespace earhost

// ---------------------------------------------------------------- C ABI
extern "C" {

int ear_wave_read(const char* path, float* out, uint64_t cap, uint64_t* n, int* channels, int* bits) {
    if (!path || !n || (cap && !out)) return -2;
    std::vector<float> s;
    int ch = 0, bps = 0;
    if (!earhost::wave_read(path, &s, &ch, &bps)) return -1;
    *n = s.size();
    if (channels) *channels = ch;
    if (bits) *bits = bps;
    if (s.size() > cap) return out ? -3 : 0;
    std::copy(s.begin(), s.end(), out);
    return 0;
}

int ear_wave_write(const char* path, const float* left, uint64_t n_left, const float* right, uint64_t n_right,
                   int norm, float norm_max) {
    if (!path || (n_left && !left) || (n_right && !right)) return -2;
    const bool ok = right ? earhost::wave_write_stereo(path, left, n_left, right, n_right, norm != 0)
                          : earhost::wave_write_mono(path, left, n_left, norm != 0, norm_max);
    return ok ? 0 : -1;
}

int ear_equalizer_split(const float* x, uint64_t n, float* low, float* mid, float* high, float f1, float f2,
                        float f3) {
    if (n && (!x || !low || !mid || !high)) return -2;
    earhost::equalizer_split(x, n, low, mid, high, f1, f2, f3);
    return 0;
}

}  // extern "C"
