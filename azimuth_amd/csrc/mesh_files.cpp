// mesh_files.cpp -- the reference importer's mesh file formats, read into
// memory for the scene builder (host code).
//
//   PLY (xml.cpp:3311-3630): "ply", "format ascii|binary_little_endian
//   <version>", comment lines, an element "vertex" of float properties and an
//   element "face" with a "list uchar|uint8 int|uint" index property (and, in
//   binary files, float properties after it, which are skipped), then
//   "end_header".  Faces of 3 indices are kept, faces of 4 are split into
//   (a, b, c) and (a, c, d), other counts are an error.  The reference reads
//   every vertex as x y z nx ny nz u v, whatever the header declares, and
//   drops the normals; here the vertex record is read by its declared float
//   properties (x, y, z and u or s, v or t where present), which is the same
//   on the files the reference reads and also reads other float layouts.
//   Vertex normals are not taken from the file (the builder computes them
//   from the faces, as the reference's post-process does).
//
//   Binary vertex file (<VertexData binaryFile=...>, xml.cpp:2397-2420): a
//   u32 count, then count x y z f32 triples.
//
//   Binary face file (<Faces binaryFile=...>, xml.cpp:2663-2700): a u32
//   count, then count u32 index triples; the element's vertexOffset is added
//   to every index and, unless the scene uses zero-based indexing, 1 is
//   subtracted.
//
// Every reader checks the file's size against what its header promises and
// returns AZ_ERR_INVALID_ARGUMENT for a malformed or truncated file,
// AZ_ERR_UNSUPPORTED for a valid PLY the reference does not read (big-endian
// or other property types).
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "azimuth_scene.h"

namespace {

bool readFile(const char *path, std::vector<char> &out) {
    if (!path) return false;
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    std::vector<char> buf;
    char chunk[1 << 16];
    size_t n;
    while ((n = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + n);
    const bool ok = !std::ferror(f);
    std::fclose(f);
    if (!ok) return false;
    buf.push_back('\0');            // parsers stop at it; it is not part of the file
    out.swap(buf);
    return true;
}

struct PlyMesh {
    std::vector<float> vertices, uvs;
    std::vector<uint32_t> faces;
};

// A cursor over the header's whitespace-separated words.
struct Words {
    const char *p, *end;
    bool next(std::string &w) {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n')) ++p;
        const char *s = p;
        while (p < end && !(*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n')) ++p;
        w.assign(s, p);
        return !w.empty();
    }
    void skipLine() {
        while (p < end && *p != '\n') ++p;
        if (p < end) ++p;
    }
};

// ASCII numbers, one after another.
struct Numbers {
    const char *p, *end;
    bool skipSpace() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n')) ++p;
        return p < end;
    }
    bool f32(float &v) {
        if (!skipSpace()) return false;
        char tmp[64];
        size_t n = 0;
        while (p + n < end && n < sizeof(tmp) - 1 && !(p[n] == ' ' || p[n] == '\t' || p[n] == '\r' || p[n] == '\n')) {
            tmp[n] = p[n];
            ++n;
        }
        tmp[n] = '\0';
        char *e = nullptr;
        v = std::strtof(tmp, &e);
        if (e == tmp || *e != '\0') return false;
        p += n;
        return true;
    }
    // (strtoull stops at the first non-digit; the file's buffer is followed
    // by a terminating zero, see readFile.)
    bool u32(uint32_t &v) {
        if (!skipSpace() || *p < '0' || *p > '9') return false;
        char *e = nullptr;
        errno = 0;
        const unsigned long long x = std::strtoull(p, &e, 10);
        if (e == p || e > end || errno != 0 || x > 0xFFFFFFFFull) return false;
        v = (uint32_t)x;
        p = e;
        return true;
    }
};

az_status parsePly(const std::vector<char> &buf, PlyMesh &out) {
    const char *data = buf.data(), *end = data + buf.size() - 1;     // the terminator excluded
    if (buf.size() < 4 || std::memcmp(data, "ply", 3) != 0) return AZ_ERR_INVALID_ARGUMENT;
    Words w{data + 3, end};
    std::string t;
    if (!w.next(t) || t != "format" || !w.next(t)) return AZ_ERR_INVALID_ARGUMENT;
    bool binary;
    if (t == "ascii") binary = false;
    else if (t == "binary_little_endian") binary = true;
    else if (t == "binary_big_endian") return AZ_ERR_UNSUPPORTED;
    else return AZ_ERR_INVALID_ARGUMENT;
    w.skipLine();                                    // the version
    uint32_t nv = 0, nf = 0;
    bool haveV = false, haveF = false;
    std::vector<std::string> vprops;                 // float properties of the vertex element
    uint32_t faceFloatsBefore = 0, faceFloatsAfter = 0;
    bool faceList = false;
    int current = 0;                                 // 1 vertex, 2 face
    for (;;) {
        if (!w.next(t)) return AZ_ERR_INVALID_ARGUMENT;
        if (t == "end_header") {
            w.skipLine();
            break;
        }
        if (t == "comment" || t == "obj_info") {
            w.skipLine();
            continue;
        }
        if (t == "element") {
            std::string name, count;
            if (!w.next(name) || !w.next(count)) return AZ_ERR_INVALID_ARGUMENT;
            char *e = nullptr;
            const unsigned long long c = std::strtoull(count.c_str(), &e, 10);
            if (*e != '\0' || c > 0xFFFFFFFFull) return AZ_ERR_INVALID_ARGUMENT;
            if (name == "vertex" && !haveV) {
                nv = (uint32_t)c;
                haveV = true;
                current = 1;
            } else if (name == "face" && !haveF) {
                nf = (uint32_t)c;
                haveF = true;
                current = 2;
            } else {
                return AZ_ERR_UNSUPPORTED;            // xml.cpp: "Unexpected element"
            }
            continue;
        }
        if (t == "property") {
            std::string type;
            if (!w.next(type)) return AZ_ERR_INVALID_ARGUMENT;
            if (type == "float" || type == "float32") {
                std::string name;
                if (!w.next(name)) return AZ_ERR_INVALID_ARGUMENT;
                if (current == 1) vprops.push_back(name);
                else if (current == 2) (faceList ? faceFloatsAfter : faceFloatsBefore)++;
                else return AZ_ERR_INVALID_ARGUMENT;
            } else if (type == "list") {
                std::string ct, it, name;
                if (!w.next(ct) || !w.next(it) || !w.next(name)) return AZ_ERR_INVALID_ARGUMENT;
                if (current != 2 || faceList) return AZ_ERR_UNSUPPORTED;
                if (!(ct == "uchar" || ct == "uint8")) return AZ_ERR_UNSUPPORTED;      // "Invalid list index"
                if (!(it == "int" || it == "uint" || it == "int32" || it == "uint32"))
                    return AZ_ERR_UNSUPPORTED;                                      // "Invalid face index"
                faceList = true;
            } else {
                return AZ_ERR_UNSUPPORTED;                                          // "Unknown property type"
            }
            continue;
        }
        return AZ_ERR_INVALID_ARGUMENT;                                             // "Expected: 'element'"
    }
    if (!haveV || (nf > 0 && !faceList)) return AZ_ERR_INVALID_ARGUMENT;
    int ix = -1, iy = -1, iz = -1, iu = -1, iv = -1;
    for (int k = 0; k < (int)vprops.size(); ++k) {
        const std::string &n = vprops[k];
        if (n == "x") ix = k;
        else if (n == "y") iy = k;
        else if (n == "z") iz = k;
        else if ((n == "u" || n == "s") && iu < 0) iu = k;
        else if ((n == "v" || n == "t") && iv < 0) iv = k;
    }
    if (ix < 0 || iy < 0 || iz < 0) return AZ_ERR_INVALID_ARGUMENT;
    const bool uv = iu >= 0 && iv >= 0;
    const size_t stride = vprops.size();
    out.vertices.resize((size_t)nv * 3);
    if (uv) out.uvs.resize((size_t)nv * 2);
    out.faces.clear();
    out.faces.reserve((size_t)nf * 3);
    std::vector<float> rec(stride);
    const char *p = w.p;
    auto keepFace = [&](const uint32_t *ids, uint32_t k) -> bool {
        for (uint32_t j = 0; j < k; ++j)
            if (ids[j] >= nv) return false;
        if (k == 3) {
            out.faces.insert(out.faces.end(), {ids[0], ids[1], ids[2]});
        } else {
            out.faces.insert(out.faces.end(), {ids[0], ids[1], ids[2], ids[0], ids[2], ids[3]});
        }
        return true;
    };
    if (binary) {
        const size_t vbytes = stride * 4;
        if ((size_t)(end - p) < (size_t)nv * vbytes) return AZ_ERR_INVALID_ARGUMENT;
        for (uint32_t i = 0; i < nv; ++i, p += vbytes) {
            std::memcpy(rec.data(), p, vbytes);
            out.vertices[3 * (size_t)i + 0] = rec[ix];
            out.vertices[3 * (size_t)i + 1] = rec[iy];
            out.vertices[3 * (size_t)i + 2] = rec[iz];
            if (uv) {
                out.uvs[2 * (size_t)i + 0] = rec[iu];
                out.uvs[2 * (size_t)i + 1] = rec[iv];
            }
        }
        for (uint32_t f = 0; f < nf; ++f) {
            if ((size_t)(end - p) < 4 * (size_t)faceFloatsBefore + 1) return AZ_ERR_INVALID_ARGUMENT;
            p += 4 * (size_t)faceFloatsBefore;
            const uint32_t k = (uint8_t)*p++;
            if (k != 3 && k != 4) return AZ_ERR_INVALID_ARGUMENT;       // "Number of edges per face must be 3 or 4"
            if ((size_t)(end - p) < 4 * (size_t)k + 4 * (size_t)faceFloatsAfter) return AZ_ERR_INVALID_ARGUMENT;
            uint32_t ids[4];
            std::memcpy(ids, p, 4 * (size_t)k);
            p += 4 * (size_t)k + 4 * (size_t)faceFloatsAfter;
            if (!keepFace(ids, k)) return AZ_ERR_INVALID_ARGUMENT;
        }
    } else {
        Numbers n{p, end};
        for (uint32_t i = 0; i < nv; ++i) {
            for (size_t k = 0; k < stride; ++k)
                if (!n.f32(rec[k])) return AZ_ERR_INVALID_ARGUMENT;
            out.vertices[3 * (size_t)i + 0] = rec[ix];
            out.vertices[3 * (size_t)i + 1] = rec[iy];
            out.vertices[3 * (size_t)i + 2] = rec[iz];
            if (uv) {
                out.uvs[2 * (size_t)i + 0] = rec[iu];
                out.uvs[2 * (size_t)i + 1] = rec[iv];
            }
        }
        for (uint32_t f = 0; f < nf; ++f) {
            float skip;
            for (uint32_t k = 0; k < faceFloatsBefore; ++k)
                if (!n.f32(skip)) return AZ_ERR_INVALID_ARGUMENT;
            uint32_t k;
            if (!n.u32(k) || (k != 3 && k != 4)) return AZ_ERR_INVALID_ARGUMENT;
            uint32_t ids[4];
            for (uint32_t j = 0; j < k; ++j)
                if (!n.u32(ids[j])) return AZ_ERR_INVALID_ARGUMENT;
            for (uint32_t j = 0; j < faceFloatsAfter; ++j)
                if (!n.f32(skip)) return AZ_ERR_INVALID_ARGUMENT;
            if (!keepFace(ids, k)) return AZ_ERR_INVALID_ARGUMENT;
        }
    }
    return AZ_OK;
}

}  // namespace

extern "C" {

az_status az_ply_read(const char *Path, f32 *Vertices, u32 *VertexCount, u32 *Faces, u32 *FaceCount, f32 *UVs,
                      u32 *UVCount) {
    if (!Path || !VertexCount || !FaceCount) return AZ_ERR_INVALID_ARGUMENT;
    std::vector<char> buf;
    if (!readFile(Path, buf)) return AZ_ERR_INVALID_ARGUMENT;
    PlyMesh m;
    const az_status s = parsePly(buf, m);
    if (s != AZ_OK) return s;
    const u32 nv = (u32)(m.vertices.size() / 3), nf = (u32)(m.faces.size() / 3), nuv = (u32)(m.uvs.size() / 2);
    // With arrays given, the counts say their capacity (in vertices, faces
    // and uv pairs) and must hold the file's.
    if (Vertices && *VertexCount < nv) return AZ_ERR_INVALID_ARGUMENT;
    if (Faces && *FaceCount < nf) return AZ_ERR_INVALID_ARGUMENT;
    if (UVs && (!UVCount || *UVCount < nuv)) return AZ_ERR_INVALID_ARGUMENT;
    if (Vertices) std::memcpy(Vertices, m.vertices.data(), m.vertices.size() * sizeof(f32));
    if (Faces) std::memcpy(Faces, m.faces.data(), m.faces.size() * sizeof(u32));
    if (UVs) std::memcpy(UVs, m.uvs.data(), m.uvs.size() * sizeof(f32));
    *VertexCount = nv;
    *FaceCount = nf;
    if (UVCount) *UVCount = nuv;
    return AZ_OK;
}

az_status az_sb_add_ply_mesh(az_scene_builder *Builder, const char *Path, const az_mesh_desc *Mesh) {
    if (!Builder || !Path || !Mesh) return AZ_ERR_INVALID_ARGUMENT;
    std::vector<char> buf;
    if (!readFile(Path, buf)) return AZ_ERR_INVALID_ARGUMENT;
    PlyMesh m;
    const az_status s = parsePly(buf, m);
    if (s != AZ_OK) return s;
    if (m.faces.empty()) return AZ_ERR_INVALID_ARGUMENT;
    az_mesh_desc d = *Mesh;
    d.Vertices = m.vertices.data();
    d.VertexCount = (u32)(m.vertices.size() / 3);
    d.Faces = m.faces.data();
    d.FaceCount = (u32)(m.faces.size() / 3);
    d.UVs = m.uvs.empty() ? nullptr : m.uvs.data();
    d.UVCount = (u32)(m.uvs.size() / 2);
    return az_sb_add_mesh(Builder, &d);
}

az_status az_read_binary_vertices(const char *Path, f32 *Vertices, u32 *Count) {
    if (!Path || !Count) return AZ_ERR_INVALID_ARGUMENT;
    std::vector<char> buf;
    if (!readFile(Path, buf) || buf.size() < 5) return AZ_ERR_INVALID_ARGUMENT;
    u32 n;
    std::memcpy(&n, buf.data(), 4);
    if ((buf.size() - 5) / 12 < n) return AZ_ERR_INVALID_ARGUMENT;
    if (Vertices) {
        if (*Count < n) return AZ_ERR_INVALID_ARGUMENT;
        std::memcpy(Vertices, buf.data() + 4, (size_t)n * 12);
    }
    *Count = n;
    return AZ_OK;
}

az_status az_read_binary_faces(const char *Path, u32 VertexOffset, b32 ZeroBasedIndexing, u32 *Faces, u32 *Count) {
    if (!Path || !Count) return AZ_ERR_INVALID_ARGUMENT;
    std::vector<char> buf;
    if (!readFile(Path, buf) || buf.size() < 5) return AZ_ERR_INVALID_ARGUMENT;
    u32 n;
    std::memcpy(&n, buf.data(), 4);
    if ((buf.size() - 5) / 12 < n) return AZ_ERR_INVALID_ARGUMENT;
    if (Faces) {
        if (*Count < n) return AZ_ERR_INVALID_ARGUMENT;
        std::memcpy(Faces, buf.data() + 4, (size_t)n * 12);
        for (size_t i = 0; i < (size_t)n * 3; ++i) {
            Faces[i] += VertexOffset;
            if (!ZeroBasedIndexing) Faces[i] -= 1u;
        }
    }
    *Count = n;
    return AZ_OK;
}

}  // extern "C"
