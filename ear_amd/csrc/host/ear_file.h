// ear_file.h -- reader for the .EAR scene format.
//
// Element encoding (src/Datatype.cpp:25-34, 96-113; writer
// blender/render_EAR/__init__.py:154-196): a 4-byte id; `int4`, `flt4`,
// `str `, `vec3` and `tri ` carry no length word, every other id is a block
// `id int32-length payload`.  Strings are NUL-padded to a multiple of 4
// (always at least one NUL); `vec3` is three `flt4`; `tri ` is three `vec3`.
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "ear_host.h"

namespace earhost {

// A SET value (int4 / flt4 / vec3 / str).
struct SettingValue {
    char type = 0;  // 'i', 'f', 'v', 's'
    int i = 0;
    float f = 0;
    Vec3 v{0, 0, 0};
    std::string s;
};

// Everything the tracer needs from one .EAR file.
struct EarFile {
    std::map<std::string, SettingValue> settings;
    std::vector<float> keys;
    float freq[3] = {0.2f, 1.0f, 2.0f};  // kHz, src/SoundFile.cpp:240-242
    std::vector<std::unique_ptr<Material>> materials;
    std::vector<std::unique_ptr<Mesh>> meshes;
    struct Source {
        std::vector<std::string> wav;
        bool has_location = false;
        Vec3 location{0, 0, 0};
        std::vector<Vec3> anim;
        float gain = 1.0f;
        float offset = 0.0f;
        std::string mesh_material;                    // emitting mesh (SSRC/3SRC `mesh` element)
        std::vector<std::array<Vec3, 3>> mesh_tris;
    };
    std::vector<Source> sources;
    struct Listener {
        bool stereo = false;
        std::string wav;
        Vec3 location{0, 0, 0};
        std::vector<Vec3> anim;
        Vec3 right_ear{1, 0, 0};
        std::vector<Vec3> right_anim;
        float head_size = 0.2f;
        Vec3 head_abs{0.1f, 0.3f, 0.9f};  // raw, as written by the exporter
    };
    std::vector<Listener> listeners;

    bool has(const std::string& k) const { return settings.count(k) > 0; }
    int get_int(const std::string& k) const;     // throws when missing (src/Settings.cpp:60-66)
    float get_float(const std::string& k) const;
    Vec3 get_vec(const std::string& k) const;
    std::string get_string(const std::string& k) const;
};

// Parses the file; throws DatatypeException on a malformed file and returns
// false when the file cannot be opened or lacks the ".EAR" magic
// (src/Datatype.cpp:81-92).
bool read_ear_file(const std::string& path, EarFile* out);

}  // namespace earhost
