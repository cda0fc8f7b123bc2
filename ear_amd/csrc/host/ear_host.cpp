// ear_host.cpp -- host C++ API over the C ABI (see ear_host.h).
#include "ear_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <sstream>

#include "audio.h"

namespace earhost {

void throw_on_error(ear_status s, const ear_ctx* ctx, const char* what) {
    if (s == EAR_OK) return;
    std::string msg = std::string(what) + ": " + ear_status_string(s);
    if (ctx) {
        const char* e = ear_ctx_last_error(ctx);
        if (e && *e) msg += std::string(" (") + e + ")";
    }
    if (s == EAR_ERR_INVALID_MATERIAL || s == EAR_ERR_INVALID_GEOMETRY) throw DatatypeException(msg);
    throw std::runtime_error(msg);
}

// ---------------------------------------------------------------- Material
Material::Material(std::string n, const float refl[3], const float* refr, const float* spec) : name(std::move(n)) {
    // src/Material.cpp:32-59: absorption starts at 1, loses (f - 1e-9) for
    // each coefficient given, must stay >= 0, and is stored as 1 - absorption
    for (int i = 0; i < 3; ++i) {
        reflection_coefficient[i] = refl[i];
        float a = 1.0f;
        a -= refl[i] - 1e-9f;
        if (refr) {
            refraction_coefficient[i] = refr[i];
            a -= refr[i] - 1e-9f;
        }
        if (a < 0.0f) throw DatatypeException("Invalid material settings");
        absorption_coefficient[i] = 1.0f - a;
        if (spec) specularity_coefficient[i] = spec[i];
    }
    has_refraction = refr != nullptr;
    has_specularity = spec != nullptr;
}

ear_material Material::to_abi() const {
    ear_material m{};
    for (int i = 0; i < 3; ++i) {
        m.refl[i] = reflection_coefficient[i];
        m.refr[i] = refraction_coefficient[i];
        m.spec[i] = specularity_coefficient[i];
    }
    m.flags = (has_refraction ? EAR_MAT_HAS_REFR : 0u) | (has_specularity ? EAR_MAT_HAS_SPEC : 0u);
    return m;
}

// ---------------------------------------------------------------- Triangle
static Vec3 vsub(const Vec3& a, const Vec3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
static Vec3 vcross(const Vec3& a, const Vec3& b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
static float vlen(const Vec3& a) { return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

Triangle::Triangle(const Vec3& a, const Vec3& b, const Vec3& c) {
    mVerts[0] = a;
    mVerts[1] = b;
    mVerts[2] = c;
    Vec3 n = vcross(vsub(b, a), vsub(c, a));
    float l = vlen(n);
    normal = l != 0.0f ? Vec3{n[0] / l, n[1] / l, n[2] / l} : n;
    area = vlen(vcross(vsub(b, a), vsub(c, b))) / 2.0f;  // src/Triangle.cpp:26-30
}

float Triangle::SignedVolume() const {
    const Vec3 &p1 = mVerts[0], &p2 = mVerts[1], &p3 = mVerts[2];
    const float v321 = p3[0] * p2[1] * p1[2];
    const float v231 = p2[0] * p3[1] * p1[2];
    const float v312 = p3[0] * p1[1] * p2[2];
    const float v132 = p1[0] * p3[1] * p2[2];
    const float v213 = p2[0] * p1[1] * p3[2];
    const float v123 = p1[0] * p2[1] * p3[2];
    return (1.0f / 6.0f) * (-v321 + v231 + v312 - v132 - v213 + v123);
}

// ---------------------------------------------------------------- Mesh
std::map<std::string, Material*> Mesh::materials;

Mesh::Mesh(Material* m, const std::vector<std::array<Vec3, 3>>& triangles) : material(m) {
    const float absorption = m ? 1.0f - m->absorption_coefficient[1] : 0.0f;  // src/Mesh.cpp:86
    for (const auto& t : triangles) {
        Triangle* tri = new Triangle(t[0], t[1], t[2]);
        tri->m = m;
        tris.push_back(tri);
        total_area += tri->area;
        total_weighted_area += tri->area * absorption;
    }
}

Mesh::~Mesh() {
    for (Triangle* t : tris) delete t;
}

void Mesh::Combine(Mesh* m) {
    for (Triangle* t : m->tris) tris.push_back(t);
    m->tris.clear();
    total_area += m->total_area;  // weighted area is not added (src/Mesh.cpp:116-123)
}

float Mesh::Volume() const {
    float v = 0.0f;
    for (const Triangle* t : tris) v += t->SignedVolume();
    return v;
}

// ---------------------------------------------------------------- animation
static std::unique_ptr<Keyframes> g_keys;
void Keyframes::Set(std::vector<float> k) {
    g_keys.reset(new Keyframes());
    g_keys->keys = std::move(k);
}
Keyframes* Keyframes::Get() { return g_keys.get(); }

template <class T>
Animated<T>::Animated(std::vector<T> f) : frames(std::move(f)) {
    Keyframes* k = Keyframes::Get();
    if (!k || k->keys.empty()) throw DatatypeException("Keyframe data not read");  // src/Animated.h:76-78
    if (frames.size() != k->keys.size()) throw DatatypeException("Keyframe count does not match");
}
template class Animated<Vec3>;

// ---------------------------------------------------------------- tracks
float& RecorderTrack::operator[](unsigned i) {
    if (i >= data.size()) data.resize((size_t)i + 44100, 0.0f);  // grows like FloatBuffer (src/Recorder.cpp:52-59)
    if (i > real_length) real_length = i;
    if (data.empty() || i < first_sample) first_sample = i;
    return data[i];
}
float RecorderTrack::Maximum() const {
    float x = 0.0f;
    for (unsigned i = first_sample; i < real_length && i < data.size(); ++i) x = std::max(x, std::fabs(data[i]));
    return x;
}
void RecorderTrack::Multiply(float f) {
    for (unsigned i = first_sample; i < real_length && i < data.size(); ++i) data[i] *= f;
}
void RecorderTrack::Power(float a) {
    for (unsigned i = first_sample; i < real_length && i < data.size(); ++i) {
        const float f = std::pow(std::fabs(data[i]), a);
        data[i] = data[i] < 0 ? f * -1.0f : f;
    }
}
void RecorderTrack::Truncate(unsigned l) {
    if (l == 0) l = 1;
    if (l >= data.size()) data.resize(l, 0.0f);
    real_length = l;
}
unsigned RecorderTrack::getLength(float tresh) const {
    if (tresh < 0.0f) return real_length;
    unsigned mx = 0;
    for (unsigned i = first_sample; i < data.size(); ++i)
        if (std::fabs(data[i]) >= tresh) mx = i;
    return mx + 1;
}
float RecorderTrack::T60() const {
    const float attenuation_gain = std::pow(10.0f, 60.0f / 20.0f);
    float min_gain = 0.0f, previous = -1.0f;
    bool inside = false;
    int last_significant = -1, direct_offset = -1;
    for (unsigned j = first_sample; j < real_length && j < data.size(); ++j) {
        const float s = data[j];
        if (inside) {
            if (s > min_gain) last_significant = (int)j;
        } else if (s < previous) {
            inside = true;
            min_gain = previous / attenuation_gain;
            direct_offset = (int)j;
        }
        previous = s;
    }
    // the reference reads uninitialised offsets when no decay is found
    if (last_significant < 0 || direct_offset < 0) return NAN;
    return (float)(last_significant - direct_offset) / 44100.0f;
}

void RecorderTrack::Normalize(float M, float MAX) { Multiply(M / (MAX < 0 ? Maximum() : MAX)); }

void RecorderTrack::reset_buffer() {
    data.assign(3 * 44100, 0.0f);
    first_sample = 3 * 44100 - 1;
    real_length = 0;
}

size_t RecorderTrack::fresh_buffer_length(unsigned first, unsigned last) {
    size_t L = 3 * 44100;
    if (first >= L) L = (size_t)first + 44100;
    while (last >= L) L += 44100;
    return L;
}

void RecorderTrack::Add(const RecorderTrack* other) {
    const unsigned len = other->getLength(0.0f);
    data.reserve((size_t)len + 44100);
    for (unsigned i = 0; i < len; ++i) (*this)[i] += other->at(i);
}

RecorderTrack* RecorderTrack::Process(ear_ctx* ctx, const SoundFile& section, const RecorderTrack* other) const {
    // The FFT form of the reference (src/Recorder.cpp:145-243): each track's
    // first getLength() samples are convolved with the (faded) section, the
    // M + N + 1 results written from the section's offset into a fresh
    // buffer; the cross-fade adds the second track's result, whose buffer
    // length also bounds the sum.  ear_convolve returns the sum of both
    // convolutions in one array.
    const unsigned M1 = real_length, M2 = other ? other->real_length : 0u, N = section.sample_length;
    std::vector<float> a(std::max(M1, 1u)), b(std::max(M2, 1u));
    for (unsigned i = 0; i < M1; ++i) a[i] = at(i);
    for (unsigned i = 0; i < M2; ++i) b[i] = other->at(i);
    const float zero = 0.0f;
    const unsigned n_out = std::max(M1, M2) + N + 1;
    std::vector<float> c(n_out);
    throw_on_error(ear_convolve(ctx, a.data(), M1, other ? b.data() : nullptr, M2, N ? section.data : &zero, N, c.data()),
                   ctx, "ear_convolve");
    const unsigned off = section.offset;
    RecorderTrack* r = new RecorderTrack();
    r->reset_buffer();
    r->data.reserve((size_t)off + n_out + 2 * 44100);
    const unsigned MN1 = M1 + N + 1;
    for (unsigned i = 0; i < MN1; ++i) (*r)[i + off] = c[i];
    if (other) {
        // the second track's result, a buffer of its own holding M2 + N + 1
        // samples from `off`, is added over that buffer's whole length
        const unsigned MN2 = M2 + N + 1;
        const size_t n = fresh_buffer_length(off, off + MN2 - 1);
        r->data.reserve(n + 44100);
        for (unsigned j = 0; j < n; ++j) (*r)[j] += (j >= off + MN1 && j < off + MN2) ? c[j - off] : 0.0f;
    }
    return r;
}

// ---------------------------------------------------------------- recorders
Recorder::~Recorder() {
    for (RecorderTrack* t : tracks) delete t;
    for (RecorderTrack* t : processed_tracks) delete t;
}
void Recorder::Process(ear_ctx* ctx, SoundFile* sf, float offset) {
    for (RecorderTrack* t : tracks) {
        std::unique_ptr<SoundFile> section = sf->Section(offset);
        processed_tracks.push_back(t->Process(ctx, *section));
    }
    is_processed = true;
}
void Recorder::Process(ear_ctx* ctx, SoundFile* sf, Recorder* r, float offset, float length) {
    for (size_t i = 0; i < tracks.size(); ++i) {
        std::unique_ptr<SoundFile> section = sf->Section(offset, length);
        processed_tracks.push_back(tracks[i]->Process(ctx, *section, r->tracks.at(i)));
    }
    is_processed = true;
}
void Recorder::Add(Recorder* r) {
    if (trackCount() != r->trackCount()) throw std::runtime_error("Recorder::Add: track counts differ");
    for (size_t i = 0; i < r->processed_tracks.size(); ++i) {
        if (i == processed_tracks.size()) {
            processed_tracks.push_back(new RecorderTrack());
            processed_tracks.back()->reset_buffer();
        }
        processed_tracks[i]->Add(r->processed_tracks[i]);
    }
    is_processed = true;
}
void Recorder::Normalize(float M) {
    const std::vector<RecorderTrack*>& ts = save_processed ? processed_tracks : tracks;
    float mx = -1e9f;
    for (const RecorderTrack* t : ts) mx = std::max(mx, t->Maximum());
    for (RecorderTrack* t : ts) t->Normalize(M, mx);
}
void Recorder::Multiply(float f) {
    for (RecorderTrack* t : tracks) t->Multiply(f);
}
void Recorder::Power(float a) {
    for (RecorderTrack* t : tracks) t->Power(a);
}
unsigned Recorder::getLength(float tresh) const {
    unsigned l = 0;
    for (const RecorderTrack* t : tracks) l = std::max(l, t->getLength(tresh));
    return l;
}
void Recorder::Truncate(int len) {
    for (RecorderTrack* t : tracks) t->Truncate((unsigned)len);
}
void Recorder::AddHistogram(int channel, const double* values, size_t n, long last_written) {
    RecorderTrack* t = tracks.at((size_t)channel);
    t->data.assign(n, 0.0f);
    size_t first = n, last = 0;
    for (size_t i = 0; i < n; ++i) {
        t->data[i] = (float)values[i];
        if (values[i] != 0.0) {
            first = std::min(first, i);
            last = i;
        }
    }
    t->first_sample = first == n ? 0 : (unsigned)first;
    t->real_length = first == n ? 0 : (unsigned)last + 1;
    if (last_written >= 0) {
        t->first_sample = first == n ? (unsigned)last_written : (unsigned)std::min<size_t>(first, (size_t)last_written);
        t->real_length = (unsigned)last_written;
    }
    has_samples = has_samples || first != n || last_written >= 0;
}

static void splat(RecorderTrack* tr, int s, float ampl, int w, bool skip_negative) {
    const float step = ampl / w;
    for (int i = 0; i < w; ++i) {
        if (!(skip_negative && i + s < 0)) (*tr)[(unsigned)(i + s)] += ampl;
        ampl -= step;
    }
}

MonoRecorder::MonoRecorder() { tracks.push_back(new RecorderTrack()); }
MonoRecorder::MonoRecorder(const Vec3& loc) : MonoRecorder() { location = loc; }
Vec3 MonoRecorder::getLocation(int i) const { return (i >= 0 && animation) ? animation->Evaluate(i) : location; }
Recorder* MonoRecorder::getBlankCopy(int) const {
    MonoRecorder* r = new MonoRecorder(location);
    r->animation = animation;
    r->filename = filename;
    r->save_processed = save_processed;
    return r;
}
void MonoRecorder::Record(const Vec3&, float a, float t, float dist, int, int) {
    // src/MonoRecorder.cpp:83-97
    const int s = (int)(t * 44100.0);
    const float width = std::sqrt(dist);
    splat(tracks[0], s, 2.0f * a / width, (int)std::ceil(width), false);
    has_samples = true;
}
bool MonoRecorder::Save(const std::string& fn, bool norm, float norm_max) {
    const RecorderTrack* t = save_processed ? processed_tracks.at(0) : tracks.at(0);
    return wave_write_mono(fn, t->data.data(), std::min<size_t>(t->getLength(), t->data.size()), norm, norm_max);
}
ear_recorder MonoRecorder::to_abi(int kf) const {
    ear_recorder r{};
    r.kind = EAR_REC_MONO;
    Vec3 p = getLocation(kf);
    std::memcpy(r.pos, p.data(), sizeof(r.pos));
    r.right_ear[0] = 1.0f;
    return r;
}

StereoRecorder::StereoRecorder() {
    tracks.push_back(new RecorderTrack());
    tracks.push_back(new RecorderTrack());
}
StereoRecorder::StereoRecorder(const Vec3& loc, const Vec3& right, float head, const Vec3& ha) : StereoRecorder() {
    location = loc;
    right_ear = right;
    head_size = head;
    for (int i = 0; i < 3; ++i) head_absorption[i] = std::max(0.0f, std::pow(1.0f - ha[i], 4.0f));
}
Vec3 StereoRecorder::getLocation(int i) const { return (i >= 0 && animation) ? animation->Evaluate(i) : location; }
Vec3 StereoRecorder::getRightEar(int i) const {
    return (i >= 0 && right_ear_animation) ? right_ear_animation->Evaluate(i) : right_ear;
}
Recorder* StereoRecorder::getBlankCopy(int) const {
    StereoRecorder* r = new StereoRecorder();
    r->location = location;
    r->right_ear = right_ear;
    r->head_size = head_size;
    r->head_absorption = head_absorption;
    r->animation = animation;
    r->right_ear_animation = right_ear_animation;
    r->filename = filename;
    r->save_processed = save_processed;
    return r;
}
void StereoRecorder::Record(const Vec3& dir, float a, float t, float dist, int band, int kf) {
    // src/StereoRecorder.cpp:97-130
    const Vec3 re = getRightEar(kf);
    const float dot = dir[0] * re[0] + dir[1] * re[1] + dir[2] * re[2];
    const float td = head_size / 343.0f;
    const int s_right = (int)((t - dot * td) * 44100.0);
    const int s_left = (int)((t + dot * td) * 44100.0);
    const float width = std::sqrt(dist);
    const float ampl = 2.0f * a / width;
    float al = ampl, ar = ampl;
    const float factor = std::pow(head_absorption[band], std::fabs(dot) * head_size);
    if (dot < 0)
        ar *= factor * factor;
    else
        al *= factor * factor;
    const int w = (int)std::ceil(width);
    splat(tracks[0], s_left, al, w, true);
    splat(tracks[1], s_right, ar, w, true);
    has_samples = true;
}
bool StereoRecorder::Save(const std::string& fn, bool norm, float) {
    const RecorderTrack* l = save_processed ? processed_tracks.at(0) : tracks.at(0);
    const RecorderTrack* r = save_processed ? processed_tracks.at(1) : tracks.at(1);
    return wave_write_stereo(fn, l->data.data(), std::min<size_t>(l->getLength(), l->data.size()), r->data.data(),
                             std::min<size_t>(r->getLength(), r->data.size()), norm);
}
ear_recorder StereoRecorder::to_abi(int kf) const {
    ear_recorder r{};
    r.kind = EAR_REC_STEREO;
    Vec3 p = getLocation(kf), e = getRightEar(kf);
    std::memcpy(r.pos, p.data(), sizeof(r.pos));
    std::memcpy(r.right_ear, e.data(), sizeof(r.right_ear));
    r.head_size = head_size;
    std::memcpy(r.head_abs, head_absorption.data(), sizeof(r.head_abs));
    return r;
}

AbstractSoundFile::~AbstractSoundFile() {
    delete mesh;
    delete animation;
}
Vec3 AbstractSoundFile::getLocation(int i) const { return (i >= 0 && animation) ? animation->Evaluate(i) : location; }

// ---------------------------------------------------------------- sound files
float SoundFile::f1 = 0.2f;  // src/SoundFile.cpp:240-242
float SoundFile::f2 = 1.0f;
float SoundFile::f3 = 2.0f;

SoundFile::SoundFile(std::vector<float> s, unsigned off) {
    SetSamples(std::move(s));
    offset = off;
}
SoundFile::SoundFile(const float* d, unsigned n, unsigned off) : data(d), sample_length(n) { offset = off; }
void SoundFile::SetSamples(std::vector<float> s) {
    samples = std::move(s);
    data = samples.data();
    sample_length = (unsigned)samples.size();
    for (auto& b : bands_) b.reset();
}
bool SoundFile::Load(const std::string& wav) {
    std::vector<float> s;
    if (!wave_read(wav, &s)) return false;
    filename = wav;
    SetSamples(std::move(s));
    return true;
}
SoundFile* SoundFile::Band(int i) {
    if (i < 0 || i > 2) throw std::out_of_range("SoundFile::Band: band index");
    if (!bands_[0]) {
        // src/SoundFile.cpp:149-160: the split signals start at offset 0
        std::vector<float> lo(sample_length), mi(sample_length), hi(sample_length);
        equalizer_split(data, sample_length, lo.data(), mi.data(), hi.data(), f1 * 1000.0f, f2 * 1000.0f,
                        f3 * 1000.0f);
        bands_[0].reset(new SoundFile(std::move(lo), 0));
        bands_[1].reset(new SoundFile(std::move(mi), 0));
        bands_[2].reset(new SoundFile(std::move(hi), 0));
    }
    return bands_[i].get();
}
std::unique_ptr<SoundFile> SoundFile::Section(unsigned start, unsigned length) const {
    // src/SoundFile.cpp:164-168
    if (start >= sample_length) return std::unique_ptr<SoundFile>(new SoundFile(nullptr, 0, 0));
    return std::unique_ptr<SoundFile>(new SoundFile(data + start, std::min(length, sample_length - start), offset + start));
}
std::unique_ptr<SoundFile> SoundFile::Section(float start, float length) const {
    // src/SoundFile.cpp:169-173: seconds at 44.1 kHz, truncated
    const unsigned s = (unsigned)(int)(start * 44100.0f);
    const unsigned l = length < 0 ? sample_length - s : (unsigned)(int)(length * 44100.0f);
    return Section(s, l);
}
int TripleBandSoundFile::Load(const std::string wav[3]) {
    for (int i = 0; i < 3; ++i) {
        std::vector<float> s;
        if (!wave_read(wav[i], &s)) return i;
        files[i].reset(new SoundFile(std::move(s), 0));  // src/SoundFile.cpp:88-92: offset 0
        filenames[i] = wav[i];
    }
    return -1;
}

// ---------------------------------------------------------------- Scene
Scene::Scene() = default;

Scene::~Scene() {
    if (dev_scene_) ear_scene_destroy(dev_scene_);
    if (ctx_) ear_ctx_destroy(ctx_);
    for (Recorder* r : listeners) delete r;
    for (AbstractSoundFile* s : sources) delete s;
    for (Mesh* m : meshes) delete m;
}

void Scene::addMesh(Mesh* m) {
    if (meshes.empty())
        meshes.push_back(m);
    else {
        meshes[0]->Combine(m);
        delete m;
    }
}

void Scene::addMaterial(Material* m) { Mesh::materials[m->name] = m; }

ear_ctx* Scene::context() {
    if (!ctx_) throw_on_error(ear_ctx_create(options.device, &ctx_), nullptr, "ear_ctx_create");
    return ctx_;
}

void Scene::ensure_device() {
    if (!ctx_) throw_on_error(ear_ctx_create(options.device, &ctx_), nullptr, "ear_ctx_create");
    const size_t n = meshes.empty() ? 0 : meshes[0]->tris.size();
    if (dev_scene_ && uploaded_tris_ == n) return;
    if (dev_scene_) {
        ear_scene_destroy(dev_scene_);
        dev_scene_ = nullptr;
    }
    std::vector<float> v0(3 * n), v1(3 * n), v2(3 * n);
    std::vector<uint16_t> mid(n);
    std::vector<ear_material> mats;
    std::map<const Material*, uint16_t> index;
    for (size_t i = 0; i < n; ++i) {
        const Triangle* t = meshes[0]->tris[i];
        for (int k = 0; k < 3; ++k) {
            v0[3 * i + k] = t->mVerts[0][k];
            v1[3 * i + k] = t->mVerts[1][k];
            v2[3 * i + k] = t->mVerts[2][k];
        }
        if (!t->m) throw DatatypeException("triangle without material");
        auto it = index.find(t->m);
        if (it == index.end()) {
            it = index.emplace(t->m, (uint16_t)mats.size()).first;
            mats.push_back(t->m->to_abi());
        }
        mid[i] = it->second;
    }
    ear_bvh_opts o{options.leaf_size, options.bvh_width, options.bvh_builder, 0};
    throw_on_error(ear_scene_upload(ctx_, v0.data(), v1.data(), v2.data(), mid.data(), (uint32_t)n,
                                    mats.empty() ? nullptr : mats.data(), (uint32_t)mats.size(), &o, &dev_scene_),
                   ctx_, "ear_scene_upload");
    uploaded_tris_ = n;
}

void Scene::Render(int band, int sound, float absorbtion_factor, int num_samples, float dry,
                   const std::vector<Recorder*>& rec, int keyframeID) {
    if (band < 0 || band > 2) throw std::runtime_error("band out of range");
    float air[3] = {1.0f, 1.0f, 1.0f};
    air[band] = absorbtion_factor;
    std::vector<Recorder*> per[3];
    per[band] = rec;
    RenderBands(sound, air, num_samples, dry, per, keyframeID, 1u << band);
}

namespace {
// The ABI description of one (sound, keyframe) render and the host buffer
// its finished tracks come back in.  The descriptor points into this object,
// so it stays where it was built.
struct PreparedRender {
    std::vector<ear_recorder> recs;
    std::vector<float> mesh_v;
    ear_render_desc d{};
    std::vector<double> value;
    std::vector<int64_t> raw, abs_raw;  // RenderOptions::multiply_quirk
    ear_hist h{};
    int nch = 0;
    bool quirk = false;
};

void prepare_render(const Scene& scene, PreparedRender& p, int sound, const float absorbtion_factor[3],
                    int num_samples, float dry, const std::vector<Recorder*> rec[3], int keyframeID,
                    unsigned band_mask, unsigned hist_bins) {
    AbstractSoundFile* src = scene.sources.at((size_t)sound);
    // the recorders of every band context describe the same listeners
    const std::vector<Recorder*>* any = nullptr;
    for (int b = 0; b < 3; ++b)
        if ((band_mask >> b) & 1u) any = &rec[b];
    if (!any || any->empty()) throw std::runtime_error("no recorders");
    for (Recorder* r : *any) {
        p.recs.push_back(r->to_abi(keyframeID));
        p.recs.back().radius = scene.options.receiver_radius;
    }
    ear_render_desc& d = p.d;
    d.band_mask = band_mask;
    for (int b = 0; b < 3; ++b) d.air[b] = absorbtion_factor[b];
    d.dry = dry;
    if (src->isMeshSource()) {
        d.source.kind = EAR_SRC_MESH;
        for (const Triangle* t : src->mesh->tris)
            for (int v = 0; v < 3; ++v)
                for (int k = 0; k < 3; ++k) p.mesh_v.push_back(t->mVerts[v][k]);
        d.source.mesh_v = p.mesh_v.data();
        d.source.mesh_n_tris = (uint32_t)src->mesh->tris.size();
    } else {
        d.source.kind = EAR_SRC_POINT;
        Vec3 pos = src->getLocation(keyframeID);
        std::memcpy(d.source.pos, pos.data(), sizeof(d.source.pos));
    }
    d.source.gain = src->getGain();
    d.recorders = p.recs.data();
    d.n_recorders = (uint32_t)p.recs.size();
    d.n_rays = (uint32_t)std::max(0, num_samples);
    d.max_bounces = scene.options.max_bounces;
    d.energy_cutoff = scene.options.energy_cutoff;
    d.seed = scene.options.seed;
    d.keyframe = keyframeID;
    d.sound = (uint32_t)sound;
    d.hist_bins = hist_bins;
    d.estimator = scene.options.estimator;
    for (const ear_recorder& r : p.recs) p.nch += r.kind == EAR_REC_STEREO ? 2 : 1;
    p.value.assign((size_t)3 * p.nch * d.hist_bins, 0.0);
    p.h.value = p.value.data();
    p.quirk = scene.options.multiply_quirk;
    if (p.quirk) {
        // the written bins are the ones whose |ramp| sum is nonzero
        d.flags |= EAR_FLAG_ABS_HIST;
        p.raw.assign((size_t)2 * 3 * p.nch * d.hist_bins, 0);
        p.abs_raw.assign(p.raw.size(), 0);
        p.h.raw = p.raw.data();
        p.h.abs_raw = p.abs_raw.data();
    }
}

// The reference's track of one (band, channel) after Scene::Render
// (RenderOptions::multiply_quirk): bins written by a path ramp are the ones
// whose exact |ramp| sum is nonzero (every ramp weight (w - i) / w is
// positive).  The highest of them keeps its paths' sum P without 1/N, plus
// the direct sound D, without gain^2 (a path is never shorter than the
// direct sound, so its ramps reach at least as far).  With no path in the
// track, the direct sound's highest bin keeps D without gain^2.  Returns that
// bin (the track's real_length), -1 if nothing was written.
long multiply_quirk(PreparedRender& p, int band, int ch) {
    const size_t nb = p.d.hist_bins, off = ((size_t)band * p.nch + ch) * nb;
    const int64_t* ab = p.abs_raw.data() + 2 * off;
    const int64_t* rw = p.raw.data() + 2 * off;
    double* v = p.value.data() + off;
    const double g2 = (double)(p.d.source.gain * p.d.source.gain);
    for (size_t j = nb; j-- > 0;) {
        if (ab[2 * j] == 0 && ab[2 * j + 1] == 0) continue;
        const long double P = ((long double)rw[2 * j + 1] * 18446744073709551616.0L +
                               (long double)(uint64_t)rw[2 * j]) /
                              (long double)std::ldexp(1.0, EAR_FRAC_BITS);
        const long double N = (long double)std::max<uint32_t>(1, p.d.n_rays);
        const long double D = g2 != 0.0 ? (long double)v[j] / g2 - P / N : 0.0L;
        v[j] = (double)(P + D);
        return (long)j;
    }
    for (size_t j = nb; j-- > 0;) {
        if (v[j] == 0.0) continue;
        if (g2 != 0.0) v[j] /= g2;
        return (long)j;
    }
    return -1;
}

// Recorder::AddHistogram of every rendered band's tracks.
void deliver_render(PreparedRender& p, const std::vector<Recorder*> rec[3], unsigned band_mask) {
    for (int b = 0; b < 3; ++b) {
        if (!((band_mask >> b) & 1u)) continue;
        int ch = 0;
        for (Recorder* r : rec[b]) {
            for (int k = 0; k < r->trackCount(); ++k, ++ch) {
                const long last = p.quirk ? multiply_quirk(p, b, ch) : -1;
                r->AddHistogram(k, p.value.data() + ((size_t)b * p.nch + ch) * p.d.hist_bins, p.d.hist_bins, last);
            }
        }
    }
}
}  // namespace

unsigned Scene::grown_bins(const ear_stats& st, unsigned bins) const {
    if (!options.grow_histogram || st.dropped_ramps == 0 || st.bins_needed <= bins) return 0;
    // whole seconds, as FloatBuffer grows (src/Recorder.cpp:52-59)
    const uint64_t need = ((uint64_t)st.bins_needed + 44099) / 44100 * 44100;
    const uint64_t g = std::min<uint64_t>(need, options.max_hist_bins);
    return g > bins ? (unsigned)g : 0u;
}

void Scene::RenderBands(int sound, const float absorbtion_factor[3], int num_samples, float dry,
                        const std::vector<Recorder*> rec[3], int keyframeID, unsigned band_mask) {
    ensure_device();
    for (unsigned bins = options.hist_bins;;) {
        PreparedRender p;
        prepare_render(*this, p, sound, absorbtion_factor, num_samples, dry, rec, keyframeID, band_mask, bins);
        throw_on_error(ear_render(ctx_, dev_scene_, &p.d, &p.h, &last_stats), ctx_, "ear_render");
        const unsigned more = grown_bins(last_stats, bins);
        if (!more) {
            deliver_render(p, rec, band_mask);
            return;
        }
        bins = more;
    }
}

void Scene::RenderBatch(const float absorbtion_factor[3], int num_samples, float dry, std::vector<Job>& jobs) {
    if (jobs.empty()) return;
    ensure_device();
    for (unsigned bins = options.hist_bins;;) {
        std::vector<std::unique_ptr<PreparedRender>> preps;
        std::vector<ear_batch_item> items;
        for (Job& j : jobs) {
            preps.push_back(std::make_unique<PreparedRender>());
            prepare_render(*this, *preps.back(), j.sound, absorbtion_factor, num_samples, dry, j.rec, j.keyframe,
                           j.band_mask, bins);
            items.push_back(ear_batch_item{dev_scene_, &preps.back()->d, &preps.back()->h});
        }
        throw_on_error(ear_render_batch(ctx_, items.data(), (uint32_t)items.size(), &last_stats), ctx_,
                       "ear_render_batch");
        const unsigned more = grown_bins(last_stats, bins);
        if (!more) {
            for (size_t k = 0; k < jobs.size(); ++k) deliver_render(*preps[k], jobs[k].rec, jobs[k].band_mask);
            return;
        }
        bins = more;
    }
}

SceneContext::SceneContext(Scene* scn, int b, int sf, int s, float ab, float dr, int kf)
    : scene(scn), band(b), soundfile_id(sf), keyframe_id(kf), samples(s), absorption(ab), dry_level(dr) {
    for (Recorder* r : scn->listeners) recorders.push_back(r->getBlankCopy(4));
}

}  // namespace earhost
