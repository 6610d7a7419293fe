// Binary little-endian PLY state files: surface, phi and psi per vertex.
//
// Writer: the byte layout of export_ply_float / export_ply_double with
// (phi, psi) (MeshIO.cpp:263-335), with one value (MeshIO.cpp:191-261) and
// with none (MeshIO.cpp:67-100): a fixed text header, then per vertex x, y, z
// and the values in one precision, then per face a uchar 3 and three uint32.
// Reader: import_ply (MeshIO.cpp:470-531, 701-903): coordinates in float or
// double, then an optional phi and psi of the same type (further vertex
// properties are skipped), faces as a uchar-counted list of uint or int that
// must hold exactly 3 indices (further face properties are skipped).  Unlike
// the reference it also checks the format line, the element counts, the file
// length and the index range, and reports failures instead of throwing.
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/bemhip.h"

namespace {

thread_local std::string io_err;

struct IoError {
    std::string msg;
};

enum class Prop { F32, F64, U8, I32, U32, I16, U16, I8, LIST_U8_U32, LIST_U8_I32, OTHER };

int prop_bytes(Prop p) {
    switch (p) {
        case Prop::F32: case Prop::I32: case Prop::U32: return 4;
        case Prop::F64: return 8;
        case Prop::U8: case Prop::I8: return 1;
        case Prop::I16: case Prop::U16: return 2;
        default: return 0;
    }
}

Prop scalar_prop(const std::string& t) {
    if (t == "float" || t == "float32") return Prop::F32;
    if (t == "double" || t == "float64") return Prop::F64;
    if (t == "uchar" || t == "uint8") return Prop::U8;
    if (t == "char" || t == "int8") return Prop::I8;
    if (t == "short" || t == "int16") return Prop::I16;
    if (t == "ushort" || t == "uint16") return Prop::U16;
    if (t == "int" || t == "int32") return Prop::I32;
    if (t == "uint" || t == "uint32") return Prop::U32;
    return Prop::OTHER;
}

struct Header {
    int64_t nv = -1, nf = -1;
    std::vector<Prop> vprops, fprops;
    std::vector<std::string> vnames;
    size_t data_offset = 0;
};

Header parse_header(std::ifstream& in) {
    Header h;
    std::string line;
    if (!std::getline(in, line) || line != "ply") throw IoError{"not a PLY file"};
    bool format_ok = false;
    std::vector<Prop>* cur = nullptr;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        std::istringstream ss(line);
        std::string w;
        ss >> w;
        if (w == "end_header") {
            h.data_offset = (size_t)in.tellg();
            if (!format_ok) throw IoError{"only binary_little_endian 1.0 PLY files are supported"};
            if (h.nv < 0 || h.nf < 0) throw IoError{"PLY header lacks the vertex or face element"};
            return h;
        } else if (w == "format") {
            std::string f, v;
            ss >> f >> v;
            format_ok = f == "binary_little_endian";
        } else if (w == "comment" || w == "obj_info" || w.empty()) {
            continue;
        } else if (w == "element") {
            std::string name;
            int64_t n = -1;
            ss >> name >> n;
            if (n < 0) throw IoError{"bad element line: " + line};
            if (name == "vertex") { h.nv = n; cur = &h.vprops; }
            else if (name == "face") {
                if (h.nv < 0) throw IoError{"the vertex element must come before the face element"};
                h.nf = n;
                cur = &h.fprops;
            }
            else throw IoError{"unsupported PLY element '" + name + "'"};
        } else if (w == "property") {
            if (!cur) throw IoError{"property before any element"};
            std::string t;
            ss >> t;
            if (t == "list") {
                std::string ct, it;
                ss >> ct >> it;
                if (ct != "uchar" && ct != "uint8") throw IoError{"polygon size must be uchar"};
                const Prop ip = scalar_prop(it);
                if (ip == Prop::U32) cur->push_back(Prop::LIST_U8_U32);
                else if (ip == Prop::I32) cur->push_back(Prop::LIST_U8_I32);
                else throw IoError{"face indices must be uint or int"};
            } else {
                const Prop p = scalar_prop(t);
                if (p == Prop::OTHER) throw IoError{"unsupported property type '" + t + "'"};
                cur->push_back(p);
                if (cur == &h.vprops) {
                    std::string name;
                    ss >> name;
                    h.vnames.push_back(name);
                }
            }
        } else {
            throw IoError{"invalid PLY header line: " + line};
        }
    }
    throw IoError{"PLY header has no end_header"};
}

// the value layout import_ply accepts: 3 coordinates, then 0..2 values, all
// float or all double (MeshIO.cpp:768-843)
struct VertexLayout {
    int prec = 0;        // 4 or 8
    int nvalues = 0;     // 0, 1 (phi) or 2 (phi, psi)
    size_t stride = 0;   // bytes per vertex record
};

VertexLayout vertex_layout(const Header& h) {
    VertexLayout L;
    for (Prop p : h.vprops) {
        if (p == Prop::LIST_U8_U32 || p == Prop::LIST_U8_I32) throw IoError{"list property on vertices"};
        L.stride += prop_bytes(p);
    }
    if (h.vprops.size() < 3) throw IoError{"vertices need x, y, z"};
    const Prop c = h.vprops[0];
    if ((c != Prop::F32 && c != Prop::F64) || h.vprops[1] != c || h.vprops[2] != c)
        throw IoError{"coordinates must be three floats or three doubles"};
    L.prec = c == Prop::F64 ? 8 : 4;
    if (h.vprops.size() > 3 && h.vprops[3] == c) {
        L.nvalues = 1;
        if (h.vprops.size() > 4 && h.vprops[4] == c) L.nvalues = 2;
    }
    return L;
}

size_t face_stride(const Header& h) {
    if (h.fprops.empty() || (h.fprops[0] != Prop::LIST_U8_U32 && h.fprops[0] != Prop::LIST_U8_I32))
        throw IoError{"the first face property must be a uchar-counted index list"};
    size_t rest = 0;
    for (size_t i = 1; i < h.fprops.size(); ++i) {
        if (h.fprops[i] == Prop::LIST_U8_U32 || h.fprops[i] == Prop::LIST_U8_I32)
            throw IoError{"more than one list property on faces"};
        rest += prop_bytes(h.fprops[i]);
    }
    return 1 + 12 + rest;
}

template <class F>
int guarded_io(F&& f) {
    try {
        f();
        io_err.clear();
        return BEMHIP_OK;
    } catch (const IoError& e) {
        io_err = e.msg;
        return BEMHIP_ERR_IO;
    } catch (const std::bad_alloc&) {
        io_err = "out of host memory";
        return BEMHIP_ERR_OOM;
    } catch (...) {
        io_err = "unexpected failure";
        return BEMHIP_ERR_IO;
    }
}

double load_real(const unsigned char* p, int prec) {
    if (prec == 8) {
        double d;
        std::memcpy(&d, p, 8);
        return d;
    }
    float f;
    std::memcpy(&f, p, 4);
    return (double)f;
}

}  // namespace

extern "C" {

const char* bemhip_io_last_error(void) { return io_err.c_str(); }

int bemhip_ply_info(const char* path, int64_t* N, int64_t* M, int32_t* nvalues, int32_t* precision) {
    return guarded_io([&] {
        if (!path || !N || !M) throw IoError{"null argument"};
        std::ifstream in(path, std::ios::binary);
        if (!in) throw IoError{std::string("cannot open ") + path};
        const Header h = parse_header(in);
        const VertexLayout L = vertex_layout(h);
        (void)face_stride(h);
        *N = h.nv;
        *M = h.nf;
        if (nvalues) *nvalues = L.nvalues;
        if (precision) *precision = L.prec;
    });
}

int bemhip_ply_read(const char* path, int64_t N, double* xyz, int64_t M, uint32_t* tri, double* phi, double* psi) {
    return guarded_io([&] {
        if (!path || (N > 0 && !xyz) || (M > 0 && !tri)) throw IoError{"null argument"};
        std::ifstream in(path, std::ios::binary);
        if (!in) throw IoError{std::string("cannot open ") + path};
        const Header h = parse_header(in);
        if (h.nv != N || h.nf != M) throw IoError{"vertex / face counts differ from the file's"};
        const VertexLayout L = vertex_layout(h);
        const size_t fs = face_stride(h);
        in.seekg(0, std::ios::end);
        const size_t total = (size_t)in.tellg();
        const size_t need = h.data_offset + L.stride * (size_t)N + fs * (size_t)M;
        if (total < need) throw IoError{"PLY file is shorter than its header says"};
        std::vector<unsigned char> buf(total - h.data_offset);
        in.seekg((std::streamoff)h.data_offset);
        in.read(reinterpret_cast<char*>(buf.data()), (std::streamsize)buf.size());
        if (!in) throw IoError{"read error"};
        const unsigned char* p = buf.data();
        for (int64_t i = 0; i < N; ++i, p += L.stride) {
            for (int d = 0; d < 3; ++d) xyz[3 * i + d] = load_real(p + d * L.prec, L.prec);
            if (phi) phi[i] = L.nvalues >= 1 ? load_real(p + 3 * L.prec, L.prec) : 0.0;
            if (psi) psi[i] = L.nvalues >= 2 ? load_real(p + 4 * L.prec, L.prec) : 0.0;
        }
        const bool sgn = h.fprops[0] == Prop::LIST_U8_I32;
        for (int64_t f = 0; f < M; ++f, p += fs) {
            if (p[0] != 3) throw IoError{"face " + std::to_string(f) + " is not a triangle"};
            for (int k = 0; k < 3; ++k) {
                uint32_t u;
                std::memcpy(&u, p + 1 + 4 * k, 4);
                const int64_t v = sgn ? (int64_t)(int32_t)u : (int64_t)u;
                if (v < 0 || v >= N) throw IoError{"face " + std::to_string(f) + " has an index out of range"};
                tri[3 * f + k] = (uint32_t)v;
            }
        }
    });
}

int bemhip_ply_write(const char* path, int64_t N, const double* xyz, int64_t M, const uint32_t* tri,
                     const double* phi, const double* psi, int32_t precision) {
    return guarded_io([&] {
        if (!path || (N > 0 && !xyz) || (M > 0 && !tri) || N < 0 || M < 0) throw IoError{"null argument"};
        if (precision != 4 && precision != 8) throw IoError{"precision must be 4 (float) or 8 (double)"};
        if (psi && !phi) throw IoError{"psi needs phi"};
        for (int64_t e = 0; e < 3 * M; ++e)
            if (tri[e] >= (uint64_t)N) throw IoError{"vertex index out of range"};
        const char* ty = precision == 8 ? "double" : "float";
        std::ostringstream hd;
        hd << "ply\nformat binary_little_endian 1.0\ncomment Output from BEM simulation\n";
        hd << "element vertex " << N << "\n";
        for (const char* c : {"x", "y", "z"}) hd << "property " << ty << " " << c << "\n";
        if (phi && psi) hd << "property " << ty << " phi\nproperty " << ty << " psi\n";
        else if (phi) hd << "property " << ty << " w\n";
        hd << "element face " << M << "\nproperty list uchar uint vertex_indices\nend_header\n";
        const int nv = 3 + (phi ? 1 : 0) + (psi ? 1 : 0);
        const size_t vs = (size_t)nv * precision;
        std::vector<unsigned char> body(vs * (size_t)N + 13 * (size_t)M);
        unsigned char* p = body.data();
        for (int64_t i = 0; i < N; ++i) {
            double v[5] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], phi ? phi[i] : 0.0, psi ? psi[i] : 0.0};
            for (int k = 0; k < nv; ++k, p += precision) {
                if (precision == 8) {
                    std::memcpy(p, &v[k], 8);
                } else {
                    const float f = (float)v[k];
                    std::memcpy(p, &f, 4);
                }
            }
        }
        for (int64_t f = 0; f < M; ++f, p += 13) {
            p[0] = 3;
            std::memcpy(p + 1, tri + 3 * f, 12);
        }
        std::ofstream out(path, std::ios::binary | std::ios::trunc);
        if (!out) throw IoError{std::string("cannot create ") + path};
        const std::string s = hd.str();
        out.write(s.data(), (std::streamsize)s.size());
        out.write(reinterpret_cast<const char*>(body.data()), (std::streamsize)body.size());
        if (!out) throw IoError{"write error"};
    });
}

}  // extern "C"
