// PLY reading and writing for the KNNCUDA command line.
//
// Reading accepts ascii, binary_little_endian and binary_big_endian files with any scalar property types;
// it takes x, y, z (required), nx, ny, nz and r, g, b / red, green, blue from the vertex element and the
// triangle list "vertex_indices" (or "vertex_index") from the face element, skipping every other element
// and property. Polygons other than triangles are rejected.
//
// Writing produces the layout of the reference's WritePLY (MeshPLY.cpp:67-127): binary little endian,
// vertex x, y, z floats, then nx, ny, nz floats when the mesh has normals, then red, green, blue uchars when
// it has colours, then faces as a uchar count 3 followed by three uint32 indices.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "egt.hpp"

namespace ply {

enum class Ty { I8, U8, I16, U16, I32, U32, F32, F64 };

inline Ty parse_type(const std::string& s) {
    if (s == "char" || s == "int8") return Ty::I8;
    if (s == "uchar" || s == "uint8") return Ty::U8;
    if (s == "short" || s == "int16") return Ty::I16;
    if (s == "ushort" || s == "uint16") return Ty::U16;
    if (s == "int" || s == "int32") return Ty::I32;
    if (s == "uint" || s == "uint32") return Ty::U32;
    if (s == "float" || s == "float32") return Ty::F32;
    if (s == "double" || s == "float64") return Ty::F64;
    throw std::runtime_error("PLY: unknown property type '" + s + "'");
}

inline int type_size(Ty t) {
    switch (t) {
        case Ty::I8:
        case Ty::U8: return 1;
        case Ty::I16:
        case Ty::U16: return 2;
        case Ty::I32:
        case Ty::U32:
        case Ty::F32: return 4;
        case Ty::F64: return 8;
    }
    return 0;
}

struct Property {
    std::string name;
    Ty type = Ty::F32;
    bool is_list = false;
    Ty count_type = Ty::U8;
};

struct Element {
    std::string name;
    uint64_t count = 0;
    std::vector<Property> props;
    int find(const char* n) const {
        for (size_t i = 0; i < props.size(); ++i)
            if (props[i].name == n) return (int)i;
        return -1;
    }
};

inline bool little_endian_host() {
    const uint16_t one = 1;
    unsigned char c;
    std::memcpy(&c, &one, 1);
    return c == 1;
}

// Decodes one binary scalar of type t at src (file byte order: swap when it differs from the host's), as a
// double.
inline double decode(const unsigned char* src, Ty t, bool swap) {
    unsigned char b[8];
    const int sz = type_size(t);
    std::memcpy(b, src, sz);
    if (swap)
        for (int i = 0; i < sz / 2; ++i) std::swap(b[i], b[sz - 1 - i]);
    switch (t) {
        case Ty::I8: return (double)(int8_t)b[0];
        case Ty::U8: return (double)b[0];
        case Ty::I16: { int16_t v; std::memcpy(&v, b, 2); return v; }
        case Ty::U16: { uint16_t v; std::memcpy(&v, b, 2); return v; }
        case Ty::I32: { int32_t v; std::memcpy(&v, b, 4); return v; }
        case Ty::U32: { uint32_t v; std::memcpy(&v, b, 4); return v; }
        case Ty::F32: { float v; std::memcpy(&v, b, 4); return v; }
        case Ty::F64: { double v; std::memcpy(&v, b, 8); return v; }
    }
    return 0.0;
}

// Reads one scalar of type t from a binary stream or an ascii token stream, as a double.
class ValueReader {
public:
    ValueReader(std::istream& in, int format) : in_(in), format_(format) {}
    double read(Ty t) {
        if (format_ == 0) {
            std::string tok;
            if (!(in_ >> tok)) throw std::runtime_error("PLY: unexpected end of ascii data");
            return std::stod(tok);
        }
        unsigned char b[8];
        const int sz = type_size(t);
        if (!in_.read(reinterpret_cast<char*>(b), sz)) throw std::runtime_error("PLY: unexpected end of binary data");
        return decode(b, t, little_endian_host() != (format_ == 1));
    }

private:
    std::istream& in_;
    int format_;  // 0 ascii, 1 binary LE, 2 binary BE
};

inline void read(const std::string& path, Mesh* mesh) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("PLY: cannot open " + path);
    std::string line;
    if (!std::getline(in, line) || line.compare(0, 3, "ply") != 0) throw std::runtime_error("PLY: missing magic");
    int format = -1;
    std::vector<Element> elems;
    for (;;) {
        if (!std::getline(in, line)) throw std::runtime_error("PLY: header without end_header");
        if (!line.empty() && line.back() == '\r') line.pop_back();
        std::istringstream ls(line);
        std::string w;
        ls >> w;
        if (w == "end_header") break;
        if (w == "format") {
            std::string f;
            ls >> f;
            format = f == "ascii" ? 0 : f == "binary_little_endian" ? 1 : f == "binary_big_endian" ? 2 : -1;
            if (format < 0) throw std::runtime_error("PLY: unknown format " + f);
        } else if (w == "element") {
            Element e;
            ls >> e.name >> e.count;
            elems.push_back(e);
        } else if (w == "property") {
            if (elems.empty()) throw std::runtime_error("PLY: property before element");
            Property p;
            std::string t;
            ls >> t;
            if (t == "list") {
                std::string ct, it;
                ls >> ct >> it >> p.name;
                p.is_list = true;
                p.count_type = parse_type(ct);
                p.type = parse_type(it);
            } else {
                p.type = parse_type(t);
                ls >> p.name;
            }
            elems.back().props.push_back(p);
        }
    }
    if (format < 0) throw std::runtime_error("PLY: missing format line");
    mesh->clear();
    ValueReader rd(in, format);
    for (const Element& e : elems) {
        const bool is_vertex = e.name == "vertex", is_face = e.name == "face";
        int ix = -1, iy = -1, iz = -1, inx = -1, iny = -1, inz = -1, ir = -1, ig = -1, ib = -1, iface = -1;
        if (is_vertex) {
            ix = e.find("x"), iy = e.find("y"), iz = e.find("z");
            if (ix < 0 || iy < 0 || iz < 0) throw std::runtime_error("PLY: vertex element without x, y, z");
            inx = e.find("nx"), iny = e.find("ny"), inz = e.find("nz");
            ir = e.find("r"), ig = e.find("g"), ib = e.find("b");
            if (ir < 0 || ig < 0 || ib < 0) ir = e.find("red"), ig = e.find("green"), ib = e.find("blue");
            mesh->mV.resize(e.count);
            if (inx >= 0 && iny >= 0 && inz >= 0) mesh->mN.resize(e.count);
            if (ir >= 0 && ig >= 0 && ib >= 0) mesh->mC.resize(e.count);
        }
        if (is_face) {
            iface = e.find("vertex_indices");
            if (iface < 0) iface = e.find("vertex_index");
            if (iface >= 0) mesh->mF.resize(e.count);
        }
        // Binary fast paths: records of fixed size are read in blocks of many records and decoded from memory
        // (elements without list properties, and a face element whose only property is the index list with a
        // one-byte count and 4-byte indices, where every face must be a triangle anyway).
        if (format != 0) {
            const bool swap = little_endian_host() != (format == 1);
            bool fixed = true;
            size_t rec = 0;
            std::vector<size_t> off(e.props.size());
            for (size_t pi = 0; pi < e.props.size(); ++pi) {
                off[pi] = rec;
                if (e.props[pi].is_list) fixed = false;
                else rec += type_size(e.props[pi].type);
            }
            const bool tri_list = iface >= 0 && e.props.size() == 1 && type_size(e.props[0].count_type) == 1 &&
                                  type_size(e.props[0].type) == 4;
            if (fixed || tri_list) {
                if (tri_list) rec = 1 + 3 * 4;
                const uint64_t block = std::max<uint64_t>(1, (64u << 20) / std::max<size_t>(rec, 1));
                std::vector<unsigned char> buf;
                for (uint64_t r0 = 0; r0 < e.count; r0 += block) {
                    const uint64_t nr = std::min<uint64_t>(block, e.count - r0);
                    buf.resize(nr * rec);
                    if (rec && !in.read(reinterpret_cast<char*>(buf.data()), (std::streamsize)(nr * rec)))
                        throw std::runtime_error("PLY: unexpected end of binary data");
                    for (uint64_t r = 0; r < nr; ++r) {
                        const unsigned char* b = buf.data() + r * rec;
                        if (tri_list) {
                            if (decode(b, e.props[0].count_type, swap) != 3.0)
                                throw std::runtime_error("PLY: only triangle faces are supported");
                            uint32_t* f = &mesh->mF[r0 + r].v0;
                            for (int q = 0; q < 3; ++q)
                                f[q] = (uint32_t)(int64_t)decode(b + 1 + 4 * q, e.props[0].type, swap);
                        } else if (is_vertex) {
                            auto v = [&](int pi) { return decode(b + off[pi], e.props[pi].type, swap); };
                            mesh->mV[r0 + r] = V3{(float)v(ix), (float)v(iy), (float)v(iz)};
                            if (!mesh->mN.empty()) mesh->mN[r0 + r] = V3{(float)v(inx), (float)v(iny), (float)v(inz)};
                            if (!mesh->mC.empty())
                                mesh->mC[r0 + r] = C4{0, (uint8_t)v(ir), (uint8_t)v(ig), (uint8_t)v(ib)};
                        }
                    }
                }
                continue;
            }
        }
        std::vector<double> vals(e.props.size());
        for (uint64_t r = 0; r < e.count; ++r) {
            for (size_t pi = 0; pi < e.props.size(); ++pi) {
                const Property& p = e.props[pi];
                if (!p.is_list) {
                    vals[pi] = rd.read(p.type);
                    continue;
                }
                const double cnt = rd.read(p.count_type);
                if (cnt < 0) throw std::runtime_error("PLY: negative list length");
                const uint64_t c = (uint64_t)cnt;
                if ((int)pi == iface && c != 3) throw std::runtime_error("PLY: only triangle faces are supported");
                for (uint64_t q = 0; q < c; ++q) {
                    const double v = rd.read(p.type);
                    if ((int)pi == iface) (&mesh->mF[r].v0)[q] = (uint32_t)(int64_t)v;
                }
            }
            if (is_vertex) {
                mesh->mV[r] = V3{(float)vals[ix], (float)vals[iy], (float)vals[iz]};
                if (!mesh->mN.empty()) mesh->mN[r] = V3{(float)vals[inx], (float)vals[iny], (float)vals[inz]};
                if (!mesh->mC.empty())
                    mesh->mC[r] = C4{0, (uint8_t)vals[ir], (uint8_t)vals[ig], (uint8_t)vals[ib]};
            }
        }
    }
    if (mesh->mV.empty()) throw std::runtime_error("PLY: no vertices");
}

inline void write(const std::string& path, Mesh* mesh) {
    std::ofstream f(path, std::ios::binary | std::ios::out | std::ios::trunc);
    if (!f) throw std::runtime_error("PLY: cannot create " + path);
    const bool normals = mesh->nN() > 0, colors = mesh->nC() > 0;
    f << "ply\n";
    f << "format binary_little_endian 1.0\n";
    f << "element vertex " << mesh->nV() << "\n";
    f << "property float x\nproperty float y\nproperty float z\n";
    if (normals) f << "property float nx\nproperty float ny\nproperty float nz\n";
    if (colors) f << "property uchar red\nproperty uchar green\nproperty uchar blue\n";
    if (mesh->nF()) {
        f << "element face " << mesh->nF() << "\n";
        f << "property list uchar uint vertex_indices\n";
    }
    f << "end_header\n";
    // records are assembled in memory and written in blocks (one stream write per block, not per value)
    const size_t vrec = sizeof(V3) + (normals ? sizeof(V3) : 0) + (colors ? sizeof(C3) : 0);
    const size_t frec = 1 + sizeof(I3);
    std::vector<char> buf;
    const size_t block = 1 << 20;
    for (size_t i0 = 0; i0 < mesh->nV(); i0 += block) {
        const size_t nr = std::min(block, mesh->nV() - i0);
        buf.resize(nr * vrec);
        char* o = buf.data();
        for (size_t i = i0; i < i0 + nr; ++i) {
            std::memcpy(o, &mesh->mV[i], sizeof(V3));
            o += sizeof(V3);
            if (normals) {
                std::memcpy(o, &mesh->mN[i], sizeof(V3));
                o += sizeof(V3);
            }
            if (colors) {
                const C3 c{mesh->mC[i].r, mesh->mC[i].g, mesh->mC[i].b};
                std::memcpy(o, &c, sizeof(C3));
                o += sizeof(C3);
            }
        }
        f.write(buf.data(), (std::streamsize)buf.size());
    }
    for (size_t i0 = 0; i0 < mesh->nF(); i0 += block) {
        const size_t nr = std::min(block, mesh->nF() - i0);
        buf.resize(nr * frec);
        char* o = buf.data();
        for (size_t i = i0; i < i0 + nr; ++i) {
            *o++ = 3;
            std::memcpy(o, &mesh->mF[i], sizeof(I3));
            o += sizeof(I3);
        }
        f.write(buf.data(), (std::streamsize)buf.size());
    }
    if (!f) throw std::runtime_error("PLY: write failed for " + path);
}

}  // namespace ply
