// ear_host.h -- host C++ API of the tracer with the reference's class names.
//
// A program written against aothms/ear's Scene / Mesh / Triangle / Material /
// Recorder / SoundFile / SceneContext (src/Scene.h, src/Mesh.h, src/Triangle.h,
// src/Material.h, src/Recorder.h, src/MonoRecorder.h, src/StereoRecorder.h,
// src/SoundFile.h, src/SceneContext.h, src/Animated.h) keeps its calls;
// Scene::Render runs on the GPU through the C ABI of include/ear.h instead of
// the per-ray CPU loop of src/Scene.cpp:111-318.  Vector types are plain
// float triples instead of gmtl's.
#pragma once

#include <array>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "ear.h"

namespace earhost {

using Vec3 = std::array<float, 3>;

// DatatypeException (src/Datatype.h:30-42): invalid input reported to main.
class DatatypeException : public std::exception {
   public:
    explicit DatatypeException(std::string m) : msg_(std::move(m)) {}
    const char* what() const noexcept override { return msg_.c_str(); }

   private:
    std::string msg_;
};

// Turns an ear_status into the reference's exception types.
void throw_on_error(ear_status s, const ear_ctx* ctx, const char* what);

enum BounceType { REFLECT, REFRACT, ABSORB };

// Material (src/Material.h:31-43, src/Material.cpp:26-72).  The constructor
// computes the retention the reference stores as absorption_coefficient and
// throws DatatypeException("Invalid material settings") when refl + refr > 1.
class Material {
   public:
    std::string name;
    float reflection_coefficient[3] = {0, 0, 0};
    float refraction_coefficient[3] = {0, 0, 0};
    float absorption_coefficient[3] = {0, 0, 0};  // retention, as in the reference
    float specularity_coefficient[3] = {0, 0, 0};
    bool has_refraction = false, has_specularity = false;
    Material(std::string n, const float refl[3], const float* refr = nullptr, const float* spec = nullptr);
    bool isTransparent() const { return has_refraction; }
    ear_material to_abi() const;
};

// Triangle (src/Triangle.h:32-43): vertices, unit normal, area, material.
class Triangle {
   public:
    Vec3 mVerts[3];
    Vec3 normal;
    float area;
    Material* m = nullptr;
    Triangle(const Vec3& a, const Vec3& b, const Vec3& c);
    float SignedVolume() const;  // src/Triangle.cpp:55-67
};

// Mesh (src/Mesh.h:41-78): one material per MESH block; Scene::addMesh merges
// every mesh into the first one (src/Mesh.cpp:116-123, including the
// reference's omission of the weighted area of merged meshes).
class Mesh {
   public:
    std::vector<Triangle*> tris;
    Material* material = nullptr;
    static std::map<std::string, Material*> materials;
    Mesh() = default;
    Mesh(Material* m, const std::vector<std::array<Vec3, 3>>& triangles);
    ~Mesh();
    void Combine(Mesh* m);
    float Area() const { return total_area; }
    float TotalAbsorption() const { return total_weighted_area; }
    float Volume() const;
    float AverageAbsorption() const { return total_weighted_area / total_area; }

   private:
    float total_area = 0, total_weighted_area = 0;
};

// Keyframes / Animated<T> (src/Animated.h:35-109): keyframe i -> frames[i].
class Keyframes {
   public:
    std::vector<float> keys;
    static void Set(std::vector<float> k);
    static Keyframes* Get();
};
template <class T>
class Animated {
   public:
    std::vector<T> frames;
    explicit Animated(std::vector<T> f);
    const T& Evaluate(int i) const { return frames.at((size_t)i); }
};

class SoundFile;

// RecorderTrack (src/Recorder.h:40-134): a float impulse-response track.
// Writes through operator[] grow the buffer as FloatBuffer does (to the
// index plus one second) and raise real_length to the largest index written.
class RecorderTrack {
   public:
    std::vector<float> data;
    unsigned first_sample = 0, real_length = 0;
    float& operator[](unsigned i);
    float at(unsigned i) const { return i < data.size() ? data[i] : 0.0f; }
    float Maximum() const;
    void Multiply(float f);
    void Normalize(float M, float MAX);  // multiplies by M / MAX (MAX < 0: by M / Maximum())
    void Power(float a);
    void Truncate(unsigned l);
    unsigned getLength(float tresh = -1.0f) const;
    float T60() const;  // src/Recorder.cpp:303-340
    // The empty FloatBuffer of the reference (src/Recorder.cpp:43-48): three
    // seconds of zeros, first_sample at its last index, real_length 0.
    void reset_buffer();
    // Length of a fresh buffer after writes at first, first + 1, ..., last.
    static size_t fresh_buffer_length(unsigned first, unsigned last);
    // src/Recorder.cpp:294-301: adds other[i] for i below other's buffer
    // length (getLength(0)), growing this buffer as needed.
    void Add(const RecorderTrack* other);
    // RecorderTrack::Process (src/Recorder.cpp:145-243): the first
    // real_length samples convolved with the section, written from the
    // section's offset into a new track; with `other` the section is faded
    // out for this track and in for `other`, and other's result is added
    // (the keyframe cross-fade).  The convolution runs on the device
    // (ear_convolve).
    RecorderTrack* Process(ear_ctx* ctx, const SoundFile& section, const RecorderTrack* other = nullptr) const;
};

// Recorder (src/Recorder.h:145-221).  Record() is the reference's per-sample
// splat (kept for CPU callers); the GPU path hands histograms over in bulk
// through AddHistogram().
class Recorder {
   public:
    std::vector<RecorderTrack*> tracks;
    std::vector<RecorderTrack*> processed_tracks;  // tracks convolved with the sound
    std::string filename;                           // the listener's output file
    bool has_samples = false, is_processed = false, save_processed = false;
    virtual ~Recorder();
    virtual int trackCount() const = 0;
    virtual Vec3 getLocation(int keyframe = -1) const = 0;
    virtual Recorder* getBlankCopy(int secs = -1) const = 0;
    virtual void Record(const Vec3& dir, float a, float t, float dist, int band, int kf) = 0;
    virtual ear_recorder to_abi(int keyframe) const = 0;
    void Multiply(float f);
    void Power(float a);
    unsigned getLength(float tresh = -1.0f) const;
    void Truncate(int len);
    // Replaces track `channel` with the finalized histogram of a render.
    // last_written >= 0: real_length = last_written, the reference's highest
    // index written (RenderOptions::multiply_quirk); otherwise one past the
    // last nonzero value.
    void AddHistogram(int channel, const double* values, size_t n, long last_written = -1);
    // src/Recorder.cpp:342-364: each track convolved with the section of sf
    // from `offset` seconds (to its end, or `length` seconds with the
    // cross-fade into r's tracks), appended to processed_tracks.
    void Process(ear_ctx* ctx, SoundFile* sf, float offset = 0.0f);
    void Process(ear_ctx* ctx, SoundFile* sf, Recorder* r, float offset, float length);
    void Add(Recorder* r);    // processed tracks summed (src/Recorder.cpp:433-442)
    void Normalize(float M);  // largest |sample| of all tracks to M (src/Recorder.cpp:444-461)
    // Writes a WAV of the processed tracks when save_processed, else of the
    // tracks (src/MonoRecorder.cpp:69-78, src/StereoRecorder.cpp:81-91).
    virtual bool Save(const std::string& fn, bool norm = true, float norm_max = -1.0f) = 0;
    bool Save() { return Save(filename, false); }
};

class MonoRecorder : public Recorder {
   public:
    Vec3 location{0, 0, 0};
    Animated<Vec3>* animation = nullptr;
    MonoRecorder();
    explicit MonoRecorder(const Vec3& loc);
    int trackCount() const override { return 1; }
    Vec3 getLocation(int i = -1) const override;
    Recorder* getBlankCopy(int secs = -1) const override;
    void Record(const Vec3& dir, float a, float t, float dist, int band, int kf) override;
    ear_recorder to_abi(int keyframe) const override;
    bool Save(const std::string& fn, bool norm = true, float norm_max = 1.0f) override;
    using Recorder::Save;
};

class StereoRecorder : public Recorder {
   public:
    Vec3 location{0, 0, 0}, right_ear{1, 0, 0};
    float head_size = 0.2f;
    Vec3 head_absorption{1, 1, 1};  // stored as max(0, (1 - h)^4) (src/StereoRecorder.cpp:53-57)
    Animated<Vec3>* animation = nullptr;
    Animated<Vec3>* right_ear_animation = nullptr;
    StereoRecorder();
    StereoRecorder(const Vec3& loc, const Vec3& right, float head, const Vec3& head_abs_raw);
    int trackCount() const override { return 2; }
    Vec3 getLocation(int i = -1) const override;
    Vec3 getRightEar(int i = -1) const;
    Recorder* getBlankCopy(int secs = -1) const override;
    void Record(const Vec3& dir, float a, float t, float dist, int band, int kf) override;
    ear_recorder to_abi(int keyframe) const override;
    bool Save(const std::string& fn, bool norm = true, float norm_max = -1.0f) override;
    using Recorder::Save;
};

// AbstractSoundFile (src/SoundFile.h:37-65): location or emitting mesh, gain,
// and the band signals the impulse responses are convolved with.
class AbstractSoundFile {
   public:
    Vec3 location{0, 0, 0};
    Animated<Vec3>* animation = nullptr;
    Mesh* mesh = nullptr;
    float gain = 1.0f;
    unsigned offset = 0;  // samples; read from the file, carried by no band (see Band)
    std::string filename;
    AbstractSoundFile() = default;
    AbstractSoundFile(const AbstractSoundFile&) = delete;
    AbstractSoundFile& operator=(const AbstractSoundFile&) = delete;
    virtual ~AbstractSoundFile();  // owns mesh and animation (src/SoundFile.cpp:71-83)
    Vec3 getLocation(int i = -1) const;
    bool isMeshSource() const { return mesh != nullptr; }
    float getGain() const { return gain; }
    virtual SoundFile* Band(int i) = 0;
};

// SoundFile (src/SoundFile.h:67-94, src/SoundFile.cpp): a mono signal at
// 44.1 kHz.  Band(i) splits it once with the three-band equalizer at the
// frequencies of SetEqBands (kHz; FREQ block, default 0.2 / 1 / 2); the band
// signals start at offset 0, as the reference's.  Section views a span of the
// samples and carries its start as offset.
class SoundFile : public AbstractSoundFile {
   public:
    std::vector<float> samples;    // owned samples (empty for a view)
    const float* data = nullptr;   // samples.data() or a span of another file's
    unsigned sample_length = 0;
    static float f1, f2, f3;
    SoundFile(const Vec3& loc, float g = 1.0f) {
        location = loc;
        gain = g;
    }
    explicit SoundFile(Mesh* emitter, float g = 1.0f) {
        mesh = emitter;
        gain = g;
    }
    SoundFile(std::vector<float> s, unsigned off);
    SoundFile(const float* d, unsigned n, unsigned off);  // a view
    SoundFile(const SoundFile&) = delete;
    SoundFile& operator=(const SoundFile&) = delete;
    void SetSamples(std::vector<float> s);
    // Reads a WAV file (WaveFile::Load + ToFloat); false when it cannot.
    bool Load(const std::string& wav);
    SoundFile* Band(int i) override;
    std::unique_ptr<SoundFile> Section(unsigned start, unsigned length) const;
    std::unique_ptr<SoundFile> Section(float start, float length = -1.0f) const;
    static void SetEqBands(float a, float b, float c) {
        f1 = a;
        f2 = b;
        f3 = c;
    }

   private:
    std::unique_ptr<SoundFile> bands_[3];
};

// TripleBandSoundFile (src/SoundFile.h:96-110): one WAV file per band, used
// as given (no equalizer).
class TripleBandSoundFile : public AbstractSoundFile {
   public:
    std::unique_ptr<SoundFile> files[3];
    std::string filenames[3];
    explicit TripleBandSoundFile(const Vec3& loc, float g = 1.0f) {
        location = loc;
        gain = g;
    }
    // Reads the three WAV files; returns the index of the first that cannot be
    // read, or -1.
    int Load(const std::string wav[3]);
    SoundFile* Band(int i) override { return files[i].get(); }
};

// Run-time parameters the reference hard-codes or reads from SET.
struct RenderOptions {
    int device = 0;
    uint64_t seed = 0xEA2;
    unsigned max_bounces = 999;   // src/Scene.cpp:143 (num_bounces < 1000)
    double energy_cutoff = 1e-8;  // src/Scene.cpp:275
    unsigned hist_bins = 3 * 44100;  // the reference's initial buffer (src/Recorder.h:36)
    // The reference's tracks grow on a write past their end (src/Recorder.cpp:52-59):
    // a render whose ramps ran past hist_bins is repeated with whole seconds of
    // room for them (same seed, same paths), up to max_hist_bins.
    bool grow_histogram = true;
    unsigned max_hist_bins = 600 * 44100;
    unsigned leaf_size = 0, bvh_width = 0, bvh_builder = 0;  // 0: library defaults
    unsigned estimator = EAR_EST_NEE;  // EAR_EST_SPHERE: receiver spheres of receiver_radius
    float receiver_radius = 0.0f;
    // FloatBuffer::Multiply (src/Recorder.cpp:85-89) stops one bin short of
    // real_length, which is the highest index written (src/Recorder.cpp:52-57),
    // so after Scene::Render (src/Scene.cpp:286-316) that bin of each track
    // holds its paths' sum without 1/N and without gain^2, and every later
    // Multiply or Power skips it.  false: every bin is scaled (real_length one
    // past the last nonzero bin); true: the reference's tracks, bin for bin.
    bool multiply_quirk = false;
};

// Scene (src/Scene.h:60-79).
class Scene {
   public:
    std::vector<Recorder*> listeners;
    std::vector<AbstractSoundFile*> sources;
    std::vector<Mesh*> meshes;
    RenderOptions options;
    Scene();
    ~Scene();
    void addListener(Recorder* l) { listeners.push_back(l); }
    void addSoundSource(AbstractSoundFile* s) { sources.push_back(s); }
    void addMesh(Mesh* m);
    void addMaterial(Material* m);
    // src/Scene.h:77: renders one (sound, keyframe, band) context into recs.
    void Render(int band, int sound, float absorbtion_factor, int num_samples, float dry,
                const std::vector<Recorder*>& rec, int keyframeID = -1);
    // All three bands of one (sound, keyframe) in one device call: rec[b] are
    // the recorders of band b's context.
    void RenderBands(int sound, const float absorbtion_factor[3], int num_samples, float dry,
                     const std::vector<Recorder*> rec[3], int keyframeID = -1, unsigned band_mask = 7);
    // One (sound, keyframe) of a batch: rec[b] are the recorders of band b's
    // context (bands outside band_mask are not rendered).
    struct Job {
        int sound = 0;
        int keyframe = -1;
        unsigned band_mask = 7;
        std::vector<Recorder*> rec[3];
    };
    // Every job in one device call (ear_render_batch).  Each job gets the
    // histograms RenderBands would give it (its rays run on their own random
    // streams); the device stays full across jobs, which the tail of a single
    // render's bounce loop leaves mostly idle.
    void RenderBatch(const float absorbtion_factor[3], int num_samples, float dry, std::vector<Job>& jobs);
    // The device context (created on first use), for the convolution stage.
    ear_ctx* context();
    ear_stats last_stats{};

   private:
    // bins for a repeated render after `st` reported cut ramps at `bins`; 0: none needed
    unsigned grown_bins(const ear_stats& st, unsigned bins) const;
    ear_ctx* ctx_ = nullptr;
    ear_scene* dev_scene_ = nullptr;
    size_t uploaded_tris_ = (size_t)-1;
    void ensure_device();
};

// SceneContext (src/SceneContext.h:26-54): one render task.
class SceneContext {
   public:
    std::vector<Recorder*> recorders;
    Scene* scene;
    int band, soundfile_id, keyframe_id, samples;
    float absorption, dry_level;
    SceneContext(Scene* scn, int b, int sf, int s, float ab, float dr, int kf = -1);
    void operator()() { scene->Render(band, soundfile_id, absorption, samples, dry_level, recorders, keyframe_id); }
};

}  // namespace earhost
