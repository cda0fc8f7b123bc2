// main.cpp -- the EAR command line over the GPU tracer.
//
//   EAR render <file>     src/EAR.cpp:55-386: parse (reading the sources'
//                         WAV files), render every (sound, keyframe, band)
//                         context, Power(0.335), threshold max/256, truncate;
//                         then convolve each response with its band of the
//                         sound (keyframes cross-faded), sum per listener,
//                         normalise to 0.8 and write the listener's WAV file.
//                         SET debugdir adds the equalizer bands, the responses
//                         (WAV and the raw FloatBuffer::Write format,
//                         src/Recorder.cpp:120-123) and the per-context
//                         results; SET noprocessing (int4 > 0) stops after
//                         the responses.
//   EAR calc T60 <file>   src/EAR.cpp:246-285, :412-423: mid band, first sound
//                         and keyframe; prints T60 of the rendered response
//                         beside the Sabine and Eyring predictions.
//   EAR test              src/EAR.cpp:424-425: returns 0.
//   EAR info <file>       parses the file and prints a summary (no device).
// Settings beyond the reference's: seed (int4), maxbounces (int4), cutoff
// (flt4), irlength (flt4, initial seconds of impulse response; a render whose
// ramps run past it is repeated with room for them, as the reference's
// buffers grow, unless growir (int4) is 0), device (int4), radius (flt4:
// receiver spheres of this radius instead of the reference's point
// connections), multiplyquirk (int4: 1 keeps each track's highest written
// bin unscaled, as the reference's FloatBuffer::Multiply does).
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>

#include "audio.h"
#include "ear_file.h"
#include "ear_host.h"

using namespace earhost;

namespace {

const char* kLoMiHi[] = {"low", "mid", "high"};

// load_audio: read the sources' WAV files, as the reference's SSRC / 3SRC
// readers do while parsing (src/SoundFile.cpp:33-42, :84-92).
Scene* build_scene(EarFile& F, bool load_audio) {
    std::unique_ptr<Scene> scene(new Scene());
    if (!F.keys.empty()) Keyframes::Set(F.keys);
    SoundFile::SetEqBands(F.freq[0], F.freq[1], F.freq[2]);
    for (auto& m : F.materials) scene->addMaterial(m.get());
    for (auto& m : F.meshes) scene->addMesh(m.release());
    for (auto& s : F.sources) {
        AbstractSoundFile* src;
        if (s.wav.size() == 3) {
            TripleBandSoundFile* t = new TripleBandSoundFile(s.location, s.gain);
            scene->addSoundSource(t);
            if (load_audio) {
                const int bad = t->Load(s.wav.data());
                if (bad >= 0) throw DatatypeException("Failed to open sound file " + s.wav[(size_t)bad]);
            }
            src = t;
        } else {
            SoundFile* sf = new SoundFile(s.location, s.gain);
            scene->addSoundSource(sf);
            sf->filename = s.wav.empty() ? "" : s.wav[0];
            if (load_audio && !sf->Load(sf->filename)) throw DatatypeException("Failed to open sound file " + sf->filename);
            src = sf;
        }
        src->offset = (unsigned)(s.offset * 44100.0f);
        if (!s.anim.empty()) src->animation = new Animated<Vec3>(s.anim);
        if (!s.mesh_tris.empty()) src->mesh = new Mesh(Mesh::materials.at(s.mesh_material), s.mesh_tris);
    }
    for (auto& l : F.listeners) {
        Recorder* r;
        if (l.stereo) {
            StereoRecorder* sr = new StereoRecorder(l.location, l.right_ear, l.head_size, l.head_abs);
            if (!l.anim.empty()) sr->animation = new Animated<Vec3>(l.anim);
            if (!l.right_anim.empty()) sr->right_ear_animation = new Animated<Vec3>(l.right_anim);
            r = sr;
        } else {
            MonoRecorder* mr = new MonoRecorder(l.location);
            if (!l.anim.empty()) mr->animation = new Animated<Vec3>(l.anim);
            r = mr;
        }
        r->filename = l.wav;
        scene->addListener(r);
    }
    RenderOptions& o = scene->options;
    if (F.has("seed")) o.seed = (uint64_t)(uint32_t)F.get_int("seed");
    if (F.has("maxbounces")) o.max_bounces = (unsigned)F.get_int("maxbounces");
    if (F.has("cutoff")) o.energy_cutoff = F.get_float("cutoff");
    if (F.has("irlength")) o.hist_bins = (unsigned)std::lround(F.get_float("irlength") * 44100.0f);
    if (F.has("growir")) o.grow_histogram = F.get_int("growir") != 0;
    if (F.has("multiplyquirk")) o.multiply_quirk = F.get_int("multiplyquirk") != 0;
    if (F.has("device")) o.device = F.get_int("device");
    if (F.has("radius")) {
        o.estimator = EAR_EST_SPHERE;
        o.receiver_radius = F.get_float("radius");
    }
    return scene.release();
}

std::string context_name(int sound, int keyframe, int band) {
    std::stringstream ss;
    ss << ".sound-" << sound;
    if (keyframe != -1) ss << ".frame-" << (keyframe < 10 ? "0" : "") << keyframe;
    ss << ".band-" << band;
    return ss.str();
}

struct Ctx {
    int sound, keyframe, band;
    std::vector<Recorder*> recorders;
};

int Render(const std::string& filename, float* calc_T60 = nullptr, float* T60_Sabine = nullptr,
           float* T60_Eyring = nullptr) {
    EarFile F;
    if (!read_ear_file(filename, &F)) {
        std::cout << "Failed to read file" << std::endl;
        return 1;
    }
    if (F.settings.empty()) {
        std::cout << "No settings block found in file" << std::endl;
        return 1;
    }
    const Vec3 absorption = F.get_vec("absorption");
    const float dry_level = F.get_float("drylevel");
    const int num_samples = F.get_int("samples") / 10;  // Release build (src/EAR.cpp:78-82)
    std::unique_ptr<Scene> scene(build_scene(F, true));
    const bool has_debugdir = F.has("debugdir");
    const std::string debugdir = has_debugdir ? F.get_string("debugdir") + "/" : "";
    if (has_debugdir) {
        // equalizer output (src/EAR.cpp:122-147)
        for (size_t sf = 0; sf < scene->sources.size(); ++sf) {
            for (int band = 0; band < 3; ++band) {
                if (calc_T60 && band != 1) continue;
                const SoundFile* b = scene->sources[sf]->Band(band);
                std::stringstream ss;
                ss << debugdir << "sound-" << sf << ".band-" << band << kLoMiHi[band] << ".wav";
                wave_write_mono(ss.str(), b->data, b->sample_length, false);
            }
            if (calc_T60) break;
        }
    }
    if (scene->sources.empty()) {
        std::cout << std::endl << "No sound sources defined" << std::endl << std::endl;
        return 1;
    }
    if (scene->listeners.empty()) {
        std::cout << std::endl << "No listeners defined" << std::endl << std::endl;
        return 1;
    }
    if (scene->meshes.empty()) {
        std::cout << std::endl << "Warning: no reflective geometry" << std::endl << std::endl;
        scene->addMesh(new Mesh());
    }
    Keyframes* keys = Keyframes::Get();
    std::cout << "Rendering..." << std::endl;

    // contexts per (sound, keyframe, band) as src/EAR.cpp:170-191, all
    // rendered in one device call (three bands per (sound, keyframe) item)
    std::vector<Ctx> ctxs;
    const float air[3] = {1.0f - absorption[0], 1.0f - absorption[1], 1.0f - absorption[2]};
    std::vector<Scene::Job> jobs;
    for (int sound = 0; sound < (int)scene->sources.size(); ++sound) {
        for (int kf = keys ? 0 : -1; kf < (keys ? (int)keys->keys.size() : 0); ++kf) {
            Scene::Job j;
            j.sound = sound;
            j.keyframe = kf;
            j.band_mask = 0;
            for (int band = 0; band < 3; ++band) {
                if (calc_T60 && band != 1) continue;
                for (Recorder* l : scene->listeners) j.rec[band].push_back(l->getBlankCopy(4));
                j.band_mask |= 1u << band;
            }
            jobs.push_back(j);
            if (calc_T60) break;
        }
        if (calc_T60) break;
    }
    scene->RenderBatch(air, num_samples, dry_level, jobs);
    for (const Scene::Job& j : jobs)
        for (int band = 0; band < 3; ++band)
            if ((j.band_mask >> band) & 1u) ctxs.push_back(Ctx{j.sound, j.keyframe, band, j.rec[band]});

    // Power(0.335) and threshold truncation (src/EAR.cpp:210-244)
    float mx = 0.0f;
    for (Ctx& c : ctxs)
        for (Recorder* r : c.recorders) {
            r->Power(0.335f);
            for (RecorderTrack* t : r->tracks) mx = std::max(mx, t->Maximum());
        }
    const float treshold = mx / 256.0f;
    for (Ctx& c : ctxs) {
        int rec_id = 0;
        for (Recorder* r : c.recorders) {
            r->Truncate((int)r->getLength(treshold));
            if (has_debugdir) {
                const std::string base = debugdir + "response-" + std::to_string(rec_id) +
                                         context_name(c.sound, c.keyframe, c.band) + kLoMiHi[c.band];
                r->Save(base + ".wav", true, mx);
                const RecorderTrack* t = r->tracks[0];
                std::ofstream f((base + ".bin").c_str(), std::ios_base::binary);
                std::vector<float> d(t->data.begin(), t->data.end());
                d.resize((size_t)t->real_length + 1, 0.0f);
                f.write((const char*)d.data(), (std::streamsize)(sizeof(float) * d.size()));
            }
            ++rec_id;
        }
    }

    const bool noprocess = F.has("noprocessing") && F.get_int("noprocessing") > 0;
    if (!noprocess && !calc_T60) {
        std::cout << std::endl << "Processing data..." << std::endl;
        ear_ctx* dev = scene->context();
        // convolution of every context's responses with its band of the sound
        // (src/EAR.cpp:297-327); a keyframe other than the last cross-fades
        // into the same band's responses at the next keyframe
        for (const Ctx& c : ctxs) {
            SoundFile* band = scene->sources[(size_t)c.sound]->Band(c.band);
            if (!keys) {
                for (Recorder* r : c.recorders) r->Process(dev, band);
                continue;
            }
            const float offset = keys->keys[(size_t)c.keyframe];
            if (c.keyframe == (int)keys->keys.size() - 1) {
                for (Recorder* r : c.recorders) r->Process(dev, band, offset);
                continue;
            }
            const Ctx* next = nullptr;
            for (const Ctx& d : ctxs)
                if (d.sound == c.sound && d.keyframe == c.keyframe + 1 && d.band == c.band) next = &d;
            const float length = keys->keys[(size_t)c.keyframe + 1] - offset;
            for (size_t k = 0; k < c.recorders.size(); ++k)
                c.recorders[k]->Process(dev, band, next->recorders.at(k), offset, length);
        }
        std::cout << "Merging result..." << std::endl;
        // src/EAR.cpp:358-386.  The reference's Truncate(getLength(1e-6)) of
        // the total acts on its unprocessed tracks only, so the saved file is
        // the whole summed buffer.
        for (size_t rec_id = 0; rec_id < scene->listeners.size(); ++rec_id) {
            std::unique_ptr<Recorder> total(scene->listeners[rec_id]->getBlankCopy());
            for (const Ctx& c : ctxs) {
                Recorder* other = c.recorders[rec_id];
                other->save_processed = true;
                if (has_debugdir)
                    other->Save(debugdir + "rec-" + std::to_string(rec_id) +
                                context_name(c.sound, c.keyframe, c.band) + ".wav");
                total->Add(other);
            }
            total->save_processed = true;
            total->Normalize(0.8f);
            if (!total->Save()) std::cout << "Failed to write " << total->filename << std::endl;
        }
    } else {
        std::cout << std::endl << "Not processing data" << std::endl;
    }
    if (calc_T60) {
        const Recorder* rec = ctxs.front().recorders.front();
        *calc_T60 = rec->tracks.front()->T60();
        const Mesh* mesh = scene->meshes[0];
        const float V = mesh->Volume(), A = mesh->TotalAbsorption(), S = mesh->Area(),
                    a = mesh->AverageAbsorption(), m = absorption[1];
        if (T60_Sabine) *T60_Sabine = 0.1611f * V / (A + 4.0f * m * V);
        if (T60_Eyring) *T60_Eyring = 0.1611f * V / (-S * std::log(1.0f - a) + 4.0f * m * V);
    }
    for (Ctx& c : ctxs)
        for (Recorder* r : c.recorders) delete r;
    return 0;
}

int Info(const std::string& filename) {
    EarFile F;
    if (!read_ear_file(filename, &F)) {
        std::cout << "Failed to read file" << std::endl;
        return 1;
    }
    size_t ntri = 0;
    for (auto& m : F.meshes) ntri += m->tris.size();
    std::cout << "materials " << F.materials.size() << "\nmeshes " << F.meshes.size() << "\ntriangles " << ntri
              << "\nsources " << F.sources.size() << "\nlisteners " << F.listeners.size() << "\nkeyframes "
              << F.keys.size() << "\nsettings " << F.settings.size() << std::endl;
    if (F.has("samples")) std::cout << "samples " << F.get_int("samples") << std::endl;
    std::unique_ptr<Scene> scene(build_scene(F, false));
    if (!scene->meshes.empty()) {
        const Mesh* m = scene->meshes[0];
        std::cout.precision(9);
        std::cout << "volume " << m->Volume() << "\narea " << m->Area() << "\ntotal_absorption "
                  << m->TotalAbsorption() << std::endl;
    }
    for (auto& l : F.listeners) std::cout << (l.stereo ? "stereo" : "mono") << " listener" << std::endl;
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    std::cout << "E.A.R acoustic ray tracer (MI355X)" << std::endl;
    for (int i = 1; i < argc; ++i) {
        const std::string cmd(argv[i]);
        const std::string arg1 = (i + 1) < argc ? std::string(argv[i + 1]) : "";
        const std::string arg2 = (i + 2) < argc ? std::string(argv[i + 2]) : "";
        if (cmd == "render" && !arg1.empty()) {
            int ret = 1;
            try {
                ret = Render(arg1);
            } catch (std::exception& e) {
                std::cout << std::endl << "Error: " << e.what() << std::endl << std::endl;
            }
            return ret;
        } else if (cmd == "calc" && arg1 == "T60" && !arg2.empty()) {
            float t60 = 0, sab = 0, ey = 0;
            int ret = 1;
            try {
                ret = Render(arg2, &t60, &sab, &ey);
                std::printf("T60_ear   : %.9fs\nT60_sabine: %.9fs\nT60_eyring: %.9fs\n", t60, sab, ey);
            } catch (std::exception& e) {
                std::cout << std::endl << "Error: " << e.what() << std::endl << std::endl;
            }
            return ret;
        } else if (cmd == "info" && !arg1.empty()) {
            try {
                return Info(arg1);
            } catch (std::exception& e) {
                std::cout << "Error: " << e.what() << std::endl;
                return 1;
            }
        } else if (cmd == "test") {
            return 0;
        }
    }
    std::cout << "Usage:" << std::endl << " EAR render <filename>" << std::endl << " EAR calc T60 <filename>" << std::endl;
    return 1;
}
