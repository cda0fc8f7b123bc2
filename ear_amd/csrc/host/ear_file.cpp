// ear_file.cpp -- .EAR reader (see ear_file.h).
#include "ear_file.h"

#include <cstring>
#include <fstream>
#include <sstream>

namespace earhost {
namespace {

struct Cursor {
    const char* p;
    const char* end;
    std::string peek() const { return (end - p >= 4) ? std::string(p, 4) : std::string(); }
    void need(size_t n) const {
        if ((size_t)(end - p) < n) throw DatatypeException("unexpected end of .EAR data");
    }
    void expect(const char* id) {
        need(4);
        if (std::memcmp(p, id, 4) != 0) {
            std::stringstream ss;
            ss << "Found '" << std::string(p, 4) << "' while expecting '" << id << "'";
            throw DatatypeException(ss.str());
        }
        p += 4;
    }
    int32_t raw_i32() {
        need(4);
        int32_t v;
        std::memcpy(&v, p, 4);
        p += 4;
        return v;
    }
    float flt() {
        expect("flt4");
        need(4);
        float v;
        std::memcpy(&v, p, 4);
        p += 4;
        return v;
    }
    int integer() {
        expect("int4");
        return raw_i32();
    }
    Vec3 vec() {
        expect("vec3");
        Vec3 v;
        v[0] = flt();
        v[1] = flt();
        v[2] = flt();
        return v;
    }
    std::string str() {
        expect("str ");
        const char* s = p;
        size_t n = 0;
        while (s + n < end && s[n]) ++n;
        if (s + n >= end) throw DatatypeException("unterminated string");
        std::string out(s, n);
        p += n + (4 - n % 4);  // NUL padding to a multiple of four, at least one
        if (p > end) throw DatatypeException("string padding past the end");
        return out;
    }
    // enters a block `id len payload`; returns the cursor of its payload
    Cursor block(const char* id) {
        expect(id);
        int32_t len = raw_i32();
        if (len < 0) throw DatatypeException("negative block length");
        need((size_t)len);
        Cursor c{p, p + len};
        p += len;
        return c;
    }
    std::vector<Vec3> anim() {
        Cursor c = block("anim");
        std::vector<Vec3> out;
        while (c.p < c.end) out.push_back(c.vec());
        return out;
    }
};

}  // namespace

int EarFile::get_int(const std::string& k) const {
    auto it = settings.find(k);
    if (it == settings.end()) throw DatatypeException("Setting '" + k + "' not found");
    if (it->second.type != 'i') throw DatatypeException("Setting '" + k + "' is not an int4");
    return it->second.i;
}
float EarFile::get_float(const std::string& k) const {
    auto it = settings.find(k);
    if (it == settings.end()) throw DatatypeException("Setting '" + k + "' not found");
    if (it->second.type == 'i') return (float)it->second.i;
    if (it->second.type != 'f') throw DatatypeException("Setting '" + k + "' is not a flt4");
    return it->second.f;
}
Vec3 EarFile::get_vec(const std::string& k) const {
    auto it = settings.find(k);
    if (it == settings.end()) throw DatatypeException("Setting '" + k + "' not found");
    if (it->second.type != 'v') throw DatatypeException("Setting '" + k + "' is not a vec3");
    return it->second.v;
}
std::string EarFile::get_string(const std::string& k) const {
    auto it = settings.find(k);
    if (it == settings.end()) throw DatatypeException("Setting '" + k + "' not found");
    if (it->second.type != 's') throw DatatypeException("Setting '" + k + "' is not a str");
    return it->second.s;
}

bool read_ear_file(const std::string& path, EarFile* F) {
    std::ifstream f(path.c_str(), std::ios_base::binary);
    if (!f.good()) return false;
    std::vector<char> buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    if (buf.size() < 4 || std::memcmp(buf.data(), ".EAR", 4) != 0) return false;
    Cursor c{buf.data() + 4, buf.data() + buf.size()};
    std::map<std::string, Material*> mats;
    while (c.p < c.end) {
        const std::string id = c.peek();
        if (id == "SET ") {
            Cursor b = c.block("SET ");
            while (b.p < b.end) {
                std::string key = b.str();
                SettingValue v;
                const std::string t = b.peek();
                if (t == "int4") {
                    v.type = 'i';
                    v.i = b.integer();
                } else if (t == "flt4") {
                    v.type = 'f';
                    v.f = b.flt();
                } else if (t == "vec3") {
                    v.type = 'v';
                    v.v = b.vec();
                } else if (t == "str ") {
                    v.type = 's';
                    v.s = b.str();
                } else {
                    throw DatatypeException("unknown setting type '" + t + "'");
                }
                F->settings[key] = v;
            }
        } else if (id == "MAT ") {
            // src/Material.cpp:26-72: name, refl x3, optional refr x3, optional spec x3
            Cursor b = c.block("MAT ");
            std::string name = b.str();
            float refl[3], refr[3], spec[3];
            for (float& x : refl) x = b.flt();
            bool has_refr = false, has_spec = false;
            if (b.peek() == "flt4") {
                for (float& x : refr) x = b.flt();
                has_refr = true;
            }
            if (b.peek() == "flt4") {
                for (float& x : spec) x = b.flt();
                has_spec = true;
            }
            F->materials.emplace_back(new Material(name, refl, has_refr ? refr : nullptr, has_spec ? spec : nullptr));
            mats[name] = F->materials.back().get();
        } else if (id == "MESH") {
            // src/Mesh.cpp:77-108: material name, then triangles
            Cursor b = c.block("MESH");
            std::string mname = b.str();
            auto it = mats.find(mname);
            if (it == mats.end()) throw DatatypeException("MESH uses material '" + mname + "' before its MAT block");
            std::vector<std::array<Vec3, 3>> tris;
            while (b.p < b.end) {
                b.expect("tri ");
                std::array<Vec3, 3> t{b.vec(), b.vec(), b.vec()};
                tris.push_back(t);
            }
            F->meshes.emplace_back(new Mesh(it->second, tris));
        } else if (id == "KEYS") {
            Cursor b = c.block("KEYS");
            F->keys.clear();
            while (b.p < b.end) F->keys.push_back(b.flt());
        } else if (id == "FREQ") {
            Cursor b = c.block("FREQ");
            for (float& x : F->freq) x = b.flt();
        } else if (id == "SSRC" || id == "3SRC") {
            // src/SoundFile.cpp:33-70 and :76-112: wav name(s), location or
            // animation, optional gain, optional offset in seconds
            Cursor b = c.block(id.c_str());
            EarFile::Source s;
            int nw = id == "3SRC" ? 3 : 1;
            for (int i = 0; i < nw; ++i) s.wav.push_back(b.str());
            if (b.peek() == "anim") {
                if (F->keys.empty()) throw DatatypeException("Keyframe data not read");
                s.anim = b.anim();
            } else if (b.peek() == "mesh") {
                // an emitting mesh: Mesh::Mesh (src/Mesh.cpp:77-90) reads a
                // material name and triangles from the block
                Cursor mb = b.block("mesh");
                s.mesh_material = mb.str();
                if (mats.find(s.mesh_material) == mats.end())
                    throw DatatypeException("mesh source uses material '" + s.mesh_material + "' before its MAT block");
                while (mb.p < mb.end) {
                    mb.expect("tri ");
                    std::array<Vec3, 3> t{mb.vec(), mb.vec(), mb.vec()};
                    s.mesh_tris.push_back(t);
                }
                if (s.mesh_tris.empty()) throw DatatypeException("mesh source without triangles");
            } else {
                s.location = b.vec();
                s.has_location = true;
            }
            if (b.peek() == "flt4") s.gain = b.flt();
            if (b.peek() == "flt4") s.offset = b.flt();
            F->sources.push_back(s);
        } else if (id == "OUT1" || id == "OUT2") {
            // src/MonoRecorder.cpp:37-55, src/StereoRecorder.cpp:33-61
            Cursor b = c.block(id.c_str());
            EarFile::Listener l;
            l.stereo = id == "OUT2";
            l.wav = b.str();
            (void)b.flt();  // read and ignored by the reference
            if (b.peek() == "anim") {
                if (F->keys.empty()) throw DatatypeException("Keyframe data not read");
                l.anim = b.anim();
            } else {
                l.location = b.vec();
            }
            if (l.stereo) {
                if (b.peek() == "anim")
                    l.right_anim = b.anim();
                else
                    l.right_ear = b.vec();
                l.head_size = b.flt();
                l.head_abs = b.vec();
            }
            F->listeners.push_back(l);
        } else if (id == "VRSN") {
            (void)c.block("VRSN");
        } else {
            // unknown block: skipped (src/EAR.cpp:115-117)
            std::string skip = id;
            (void)c.block(skip.c_str());
        }
    }
    return true;
}

}  // namespace earhost
