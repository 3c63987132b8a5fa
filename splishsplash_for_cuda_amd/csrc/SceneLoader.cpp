// SPlisHSPlasH scene files and text particle files (see SceneLoader.h).
#include "SceneLoader.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <unordered_set>

namespace SPH {

// ---------------------------------------------------------------- JSON
const JsonValue* JsonValue::get(const std::string& key) const
{
    if (kind != Object) return nullptr;
    auto it = obj.find(key);
    return it == obj.end() ? nullptr : &it->second;
}

namespace {

struct JsonParser {
    const std::string& s;
    size_t i = 0;

    [[noreturn]] void fail(const char* what) const
    {
        throw std::runtime_error(std::string("JSON: ") + what + " at offset " + std::to_string(i));
    }
    void ws()
    {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) ++i;
    }
    bool lit(const char* w)
    {
        const size_t n = std::strlen(w);
        if (s.compare(i, n, w) != 0) return false;
        i += n;
        return true;
    }
    static void utf8(std::string& out, unsigned cp)
    {
        if (cp < 0x80) {
            out += (char)cp;
        } else if (cp < 0x800) {
            out += (char)(0xC0 | (cp >> 6));
            out += (char)(0x80 | (cp & 0x3F));
        } else {
            out += (char)(0xE0 | (cp >> 12));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        }
    }
    std::string str()
    {
        if (s[i] != '"') fail("expected a string");
        ++i;
        std::string out;
        while (true) {
            if (i >= s.size()) fail("unterminated string");
            const char c = s[i++];
            if (c == '"') break;
            if (c != '\\') { out += c; continue; }
            if (i >= s.size()) fail("unterminated escape");
            const char e = s[i++];
            switch (e) {
            case '"': out += '"'; break;
            case '\\': out += '\\'; break;
            case '/': out += '/'; break;
            case 'b': out += '\b'; break;
            case 'f': out += '\f'; break;
            case 'n': out += '\n'; break;
            case 'r': out += '\r'; break;
            case 't': out += '\t'; break;
            case 'u': {
                if (i + 4 > s.size()) fail("short \\u escape");
                utf8(out, (unsigned)std::strtoul(s.substr(i, 4).c_str(), nullptr, 16));
                i += 4;
                break;
            }
            default: fail("bad escape");
            }
        }
        return out;
    }
    JsonValue value()
    {
        ws();
        if (i >= s.size()) fail("unexpected end");
        JsonValue v;
        const char c = s[i];
        if (c == '{') {
            v.kind = JsonValue::Object;
            ++i;
            ws();
            if (i < s.size() && s[i] == '}') { ++i; return v; }
            while (true) {
                ws();
                const std::string k = str();
                ws();
                if (i >= s.size() || s[i] != ':') fail("expected ':'");
                ++i;
                v.obj[k] = value();
                ws();
                if (i < s.size() && s[i] == ',') { ++i; continue; }
                if (i < s.size() && s[i] == '}') { ++i; break; }
                fail("expected ',' or '}'");
            }
        } else if (c == '[') {
            v.kind = JsonValue::Array;
            ++i;
            ws();
            if (i < s.size() && s[i] == ']') { ++i; return v; }
            while (true) {
                v.arr.push_back(value());
                ws();
                if (i < s.size() && s[i] == ',') { ++i; continue; }
                if (i < s.size() && s[i] == ']') { ++i; break; }
                fail("expected ',' or ']'");
            }
        } else if (c == '"') {
            v.kind = JsonValue::String;
            v.str = str();
        } else if (lit("true")) {
            v.kind = JsonValue::Bool;
            v.b = true;
        } else if (lit("false")) {
            v.kind = JsonValue::Bool;
        } else if (lit("null")) {
            v.kind = JsonValue::Null;
        } else {
            const char* b = s.c_str() + i;
            char* e = nullptr;
            v.num = std::strtod(b, &e);
            if (e == b) fail("unexpected character");
            v.kind = JsonValue::Number;
            i += (size_t)(e - b);
        }
        return v;
    }
};

bool readNum(const JsonValue* j, double& v)
{
    if (!j) return false;
    if (j->kind == JsonValue::Number) { v = j->num; return true; }
    if (j->kind == JsonValue::Bool) { v = j->b ? 1.0 : 0.0; return true; }
    return false;
}

template <class T>
void readUint(const JsonValue* j, T& v)
{
    double d;
    if (readNum(j, d)) v = (T)d;
}

void readBool(const JsonValue* j, bool& v)
{
    double d;
    if (readNum(j, d)) v = d != 0.0;
}

void readStr(const JsonValue* j, std::string& v)
{
    if (j && j->kind == JsonValue::String) v = j->str;
}

bool readVec(const JsonValue* j, Vec3& v)
{
    if (!j || j->kind != JsonValue::Array || j->arr.size() < 3) return false;
    for (int a = 0; a < 3; ++a)
        if (!readNum(&j->arr[a], v[a])) return false;
    return true;
}

// Rotation of `angle` about `axis` (normalised), as Eigen::AngleAxis builds it.
Mat3 angleAxis(double angle, Vec3 a)
{
    const double n = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    Mat3 R{{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
    if (n == 0.0) return R;
    for (double& x : a) x /= n;
    const double c = std::cos(angle), s = std::sin(angle), t = 1.0 - c;
    R[0] = {t * a[0] * a[0] + c, t * a[0] * a[1] - s * a[2], t * a[0] * a[2] + s * a[1]};
    R[1] = {t * a[0] * a[1] + s * a[2], t * a[1] * a[1] + c, t * a[1] * a[2] - s * a[0]};
    R[2] = {t * a[0] * a[2] - s * a[1], t * a[1] * a[2] + s * a[0], t * a[2] * a[2] + c};
    return R;
}

void readPlacement(const JsonValue& m, Vec3& translation, Mat3& rotation)
{
    readVec(m.get("translation"), translation);
    Vec3 axis{0, 0, 0};
    double angle = 0.0;
    if (readVec(m.get("rotationAxis"), axis) && readNum(m.get("rotationAngle"), angle)) rotation = angleAxis(angle, axis);
}

std::string readFile(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

}  // namespace

JsonValue parse_json(const std::string& text)
{
    JsonParser p{text};
    JsonValue v = p.value();
    p.ws();
    if (p.i != text.size()) p.fail("trailing characters");
    return v;
}

// ---------------------------------------------------------------- scene
void SceneLoader::readScene(const std::string& fileName, Scene& scene)
{
    const JsonValue j = parse_json(readFile(fileName));
    if (const JsonValue* c = j.get("Configuration")) {
        readNum(c->get("timeStepSize"), scene.timeStepSize);
        readNum(c->get("particleRadius"), scene.particleRadius);
        readUint(c->get("cflMethod"), scene.cflMethod);
        readNum(c->get("cflFactor"), scene.cflFactor);
        readNum(c->get("cflMaxTimeStepSize"), scene.cflMaxTimeStepSize);
        readUint(c->get("maxIterations"), scene.maxIterations);
        readNum(c->get("maxError"), scene.maxError);
        readUint(c->get("maxIterationsV"), scene.maxIterationsV);
        readNum(c->get("maxErrorV"), scene.maxErrorV);
        readNum(c->get("viscosity"), scene.viscosity);
        readNum(c->get("density0"), scene.density0);
        readVec(c->get("gravitation"), scene.gravitation);
        readBool(c->get("enableDivergenceSolver"), scene.enableDivergenceSolver);
    }
    if (const JsonValue* rb = j.get("RigidBodies")) {
        for (const JsonValue& m : rb->arr) {
            BoundaryData d;
            readStr(m.get("geometryFile"), d.meshFile);
            readStr(m.get("particleFile"), d.samplesFile);
            if (d.meshFile.empty() && d.samplesFile.empty()) continue;
            readPlacement(m, d.translation, d.rotation);
            readVec(m.get("scale"), d.scale);
            readBool(m.get("isDynamic"), d.dynamic);
            readBool(m.get("isWall"), d.isWall);
            readNum(m.get("density"), d.density);
            readVec(m.get("velocity"), d.velocity);
            readVec(m.get("angularVelocity"), d.angularVelocity);
            scene.boundaryModels.push_back(d);
        }
    }
    if (const JsonValue* fm = j.get("FluidModels")) {
        for (const JsonValue& m : fm->arr) {
            FluidData d;
            readStr(m.get("particleFile"), d.samplesFile);
            if (d.samplesFile.empty()) continue;
            readPlacement(m, d.translation, d.rotation);
            readNum(m.get("scale"), d.scale);
            scene.fluidModels.push_back(d);
        }
    }
    if (const JsonValue* fb = j.get("FluidBlocks")) {
        for (const JsonValue& m : fb->arr) {
            Vec3 translation{0, 0, 0}, scale{1, 1, 1}, minX, maxX;
            readVec(m.get("translation"), translation);
            readVec(m.get("scale"), scale);
            if (!readVec(m.get("start"), minX) || !readVec(m.get("end"), maxX)) continue;
            FluidBlock b;
            for (int a = 0; a < 3; ++a) {
                b.box.minX[a] = scale[a] * minX[a] + translation[a];
                b.box.maxX[a] = scale[a] * maxX[a] + translation[a];
            }
            readUint(m.get("denseMode"), b.mode);
            readVec(m.get("initialVelocity"), b.initialVelocity);
            scene.fluidBlocks.push_back(b);
        }
    }
}

void createFluidBlocks(const SceneLoader::Scene& scene, std::vector<Vec3>& pos, std::vector<Vec3>& vel)
{
    const double r = scene.particleRadius, diam = 2.0 * r, eps = 1.0e-9;
    for (const SceneLoader::FluidBlock& fb : scene.fluidBlocks) {
        double xshift = diam, yshift = diam;
        if (fb.mode == 1) {
            yshift = std::sqrt(3.0) * r + eps;
        } else if (fb.mode == 2) {
            xshift = std::sqrt(6.0) * diam / 3.0 + eps;
            yshift = std::sqrt(3.0) * r + eps;
        }
        Vec3 diff{fb.box.maxX[0] - fb.box.minX[0], fb.box.maxX[1] - fb.box.minX[1], fb.box.maxX[2] - fb.box.minX[2]};
        if (fb.mode == 1) {
            diff[0] -= diam;
            diff[2] -= diam;
        } else if (fb.mode == 2) {
            diff[0] -= xshift;
            diff[2] -= diam;
        }
        const int sx = (int)std::round(diff[0] / xshift) - 1;
        const int sy = (int)std::round(diff[1] / yshift) - 1;
        const int sz = (int)std::round(diff[2] / diam) - 1;
        const Vec3 start{fb.box.minX[0] + diam, fb.box.minX[1] + diam, fb.box.minX[2] + diam};
        for (int j = 0; j < sx; ++j)
            for (int k = 0; k < sy; ++k)
                for (int l = 0; l < sz; ++l) {
                    Vec3 p{j * xshift + start[0], k * yshift + start[1], l * diam + start[2]};
                    if (fb.mode == 1) {
                        if (k % 2 == 0) p[2] += r; else p[0] += r;
                    } else if (fb.mode == 2) {
                        p[2] += r;
                        if (j % 2) p[2] += diam / (2.0 * (k % 2 ? -1 : 1));
                        if (k % 2 == 0) p[0] += xshift / 2.0;
                    }
                    pos.push_back(p);
                    vel.push_back(fb.initialVelocity);
                }
    }
}

namespace {

struct Mesh {
    std::vector<Vec3> v;
    std::vector<std::array<uint32_t, 3>> f;
};

// OBJ: vertices scaled per axis (OBJLoader::loadObj with a scale), faces as
// triangle fans; texture / normal indices are ignored.
Mesh loadObj(const std::string& path, const Vec3& scale)
{
    std::istringstream in(readFile(path));
    Mesh m;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string tag;
        ls >> tag;
        if (tag == "v") {
            Vec3 p{0, 0, 0};
            ls >> p[0] >> p[1] >> p[2];
            for (int a = 0; a < 3; ++a) p[a] *= scale[a];
            m.v.push_back(p);
        } else if (tag == "f") {
            std::vector<uint32_t> idx;
            std::string tok;
            while (ls >> tok) {
                const long k = std::strtol(tok.c_str(), nullptr, 10);
                const long id = k < 0 ? (long)m.v.size() + k : k - 1;
                if (id < 0 || id >= (long)m.v.size()) throw std::runtime_error("OBJ face index out of range: " + path);
                idx.push_back((uint32_t)id);
            }
            for (size_t t = 2; t < idx.size(); ++t) m.f.push_back({idx[0], idx[t - 1], idx[t]});
        }
    }
    if (m.v.empty()) throw std::runtime_error("OBJ without vertices: " + path);
    return m;
}

struct PointSet {
    double q;
    std::unordered_set<std::string> seen;
    std::vector<Vec3>& out;
    void add(const Vec3& p)
    {
        char key[96];
        std::snprintf(key, sizeof key, "%lld,%lld,%lld", (long long)std::llround(p[0] / q),
                      (long long)std::llround(p[1] / q), (long long)std::llround(p[2] / q));
        if (seen.insert(key).second) out.push_back(p);
    }
};

}  // namespace

namespace {

// A rigid body's mesh placed as the demos place it: scaled per axis, then
// rotated and translated (Demos/StaticBoundaryDemo/main.cpp:421-424).
Mesh placedMesh(const SceneLoader::BoundaryData& bd, const std::string& base_path)
{
    if (!bd.samplesFile.empty()) throw std::runtime_error("particle sample files are not read: " + bd.samplesFile);
    const std::string path = (!bd.meshFile.empty() && bd.meshFile[0] == '/') ? bd.meshFile : base_path + "/" + bd.meshFile;
    Mesh m = loadObj(path, bd.scale);
    for (Vec3& p : m.v) {
        Vec3 q{0, 0, 0};
        for (int a = 0; a < 3; ++a)
            q[a] = bd.rotation[a][0] * p[0] + bd.rotation[a][1] * p[1] + bd.rotation[a][2] * p[2] + bd.translation[a];
        p = q;
    }
    return m;
}

// Surface particles of a mesh at spacing d: an axis-aligned box (8 corners)
// gets the regular lattice of its six faces, any other mesh a lattice of
// spacing <= d on every triangle (coincident points kept once).
void sampleMesh(const Mesh& m, double r, std::vector<Vec3>& pos)
{
    const double d = 2.0 * r;
    PointSet ps{0.25 * r, {}, pos};
    Vec3 lo = m.v[0], hi = m.v[0];
    for (const Vec3& p : m.v)
        for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], p[a]); hi[a] = std::max(hi[a], p[a]); }
    bool box = m.v.size() == 8;
    const double tol = 1.0e-6 * (1.0 + std::fabs(hi[0] - lo[0]) + std::fabs(hi[1] - lo[1]) + std::fabs(hi[2] - lo[2]));
    for (const Vec3& p : m.v)
        for (int a = 0; a < 3; ++a)
            box = box && (std::fabs(p[a] - lo[a]) < tol || std::fabs(p[a] - hi[a]) < tol);
    if (box) {
        int n[3];
        for (int a = 0; a < 3; ++a) n[a] = std::max(0, (int)std::lround((hi[a] - lo[a]) / d));
        for (int i = 0; i <= n[0]; ++i)
            for (int j = 0; j <= n[1]; ++j)
                for (int k = 0; k <= n[2]; ++k) {
                    if (i != 0 && i != n[0] && j != 0 && j != n[1] && k != 0 && k != n[2]) continue;
                    const int c[3] = {i, j, k};
                    Vec3 p;
                    for (int a = 0; a < 3; ++a) p[a] = n[a] ? lo[a] + (hi[a] - lo[a]) * c[a] / n[a] : lo[a];
                    ps.add(p);
                }
        return;
    }
    for (const auto& t : m.f) {
        const Vec3 &a = m.v[t[0]], &b = m.v[t[1]], &c = m.v[t[2]];
        Vec3 e1, e2;
        for (int x = 0; x < 3; ++x) { e1[x] = b[x] - a[x]; e2[x] = c[x] - a[x]; }
        const int n1 = std::max(1, (int)std::ceil(std::sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]) / d));
        const int n2 = std::max(1, (int)std::ceil(std::sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]) / d));
        for (int i = 0; i <= n1; ++i)
            for (int j = 0; j <= n2; ++j) {
                const double u = (double)i / n1, w = (double)j / n2;
                if (u + w > 1.0 + 1e-12) continue;
                ps.add(Vec3{a[0] + u * e1[0] + w * e2[0], a[1] + u * e1[1] + w * e2[1], a[2] + u * e1[2] + w * e2[2]});
            }
    }
}

// Eigen decomposition of a symmetric 3x3 matrix by cyclic Jacobi rotations:
// A = E diag(l) E^T, E's columns the eigenvectors.
void symEigen(Mat3 A, Vec3& l, Mat3& E)
{
    E = Mat3{{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
    for (int sweep = 0; sweep < 50; ++sweep) {
        const double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
        const double tr = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
        if (off <= 1e-30 * tr || off == 0.0) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (A[p][q] == 0.0) continue;
                const double th = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; ++k) {       // A <- A J
                    const double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) {       // A <- J^T A
                    const double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; ++k) {       // E <- E J
                    const double ekp = E[k][p], ekq = E[k][q];
                    E[k][p] = c * ekp - s * ekq;
                    E[k][q] = s * ekp + c * ekq;
                }
            }
    }
    for (int a = 0; a < 3; ++a) l[a] = A[a][a];
}

// unit quaternion (x, y, z, w) of a rotation matrix (Shepperd's method)
std::array<double, 4> matQuat(const Mat3& R)
{
    const double tr = R[0][0] + R[1][1] + R[2][2];
    std::array<double, 4> q;
    if (tr > 0.0) {
        const double s = 2.0 * std::sqrt(tr + 1.0);
        q = {(R[2][1] - R[1][2]) / s, (R[0][2] - R[2][0]) / s, (R[1][0] - R[0][1]) / s, 0.25 * s};
    } else if (R[0][0] > R[1][1] && R[0][0] > R[2][2]) {
        const double s = 2.0 * std::sqrt(1.0 + R[0][0] - R[1][1] - R[2][2]);
        q = {0.25 * s, (R[0][1] + R[1][0]) / s, (R[0][2] + R[2][0]) / s, (R[2][1] - R[1][2]) / s};
    } else if (R[1][1] > R[2][2]) {
        const double s = 2.0 * std::sqrt(1.0 + R[1][1] - R[0][0] - R[2][2]);
        q = {(R[0][1] + R[1][0]) / s, 0.25 * s, (R[1][2] + R[2][1]) / s, (R[0][2] - R[2][0]) / s};
    } else {
        const double s = 2.0 * std::sqrt(1.0 + R[2][2] - R[0][0] - R[1][1]);
        q = {(R[0][2] + R[2][0]) / s, (R[1][2] + R[2][1]) / s, 0.25 * s, (R[1][0] - R[0][1]) / s};
    }
    const double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    for (double& x : q) x /= n;
    return q;
}

}  // namespace

void meshMassProperties(const std::vector<Vec3>& v, const std::vector<std::array<uint32_t, 3>>& f, double density,
                        double& mass, Vec3& com, Vec3& moments, Mat3& axes)
{
    // signed tetrahedra (origin, a, b, c): volume a.(b x c) / 6, first moment
    // V (a + b + c) / 4, second moments V / 20 (sum_k v_k v_k^T + s s^T), s = a + b + c
    double V = 0.0;
    Vec3 S{0, 0, 0};
    Mat3 C{};
    for (const auto& t : f) {
        const Vec3 &a = v[t[0]], &b = v[t[1]], &c = v[t[2]];
        const double vol = (a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                            a[2] * (b[0] * c[1] - b[1] * c[0])) / 6.0;
        const Vec3 s{a[0] + b[0] + c[0], a[1] + b[1] + c[1], a[2] + b[2] + c[2]};
        V += vol;
        for (int i = 0; i < 3; ++i) {
            S[i] += vol * s[i] / 4.0;
            for (int j = 0; j < 3; ++j) C[i][j] += vol / 20.0 * (a[i] * a[j] + b[i] * b[j] + c[i] * c[j] + s[i] * s[j]);
        }
    }
    if (V < 0.0) {                          // inward-facing triangles
        V = -V;
        for (int i = 0; i < 3; ++i) {
            S[i] = -S[i];
            for (int j = 0; j < 3; ++j) C[i][j] = -C[i][j];
        }
    }
    if (!(V > 0.0)) throw std::runtime_error("a dynamic body needs a closed mesh");
    mass = density * V;
    for (int i = 0; i < 3; ++i) com[i] = S[i] / V;
    // covariance about the center, then the inertia tensor tr(C) I - C
    Mat3 Cc;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Cc[i][j] = density * (C[i][j] - V * com[i] * com[j]);
    const double tr = Cc[0][0] + Cc[1][1] + Cc[2][2];
    Mat3 I;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) I[i][j] = (i == j ? tr : 0.0) - Cc[i][j];
    symEigen(I, moments, axes);
    // a proper rotation (det +1)
    const double det = axes[0][0] * (axes[1][1] * axes[2][2] - axes[1][2] * axes[2][1]) -
                       axes[0][1] * (axes[1][0] * axes[2][2] - axes[1][2] * axes[2][0]) +
                       axes[0][2] * (axes[1][0] * axes[2][1] - axes[1][1] * axes[2][0]);
    if (det < 0.0)
        for (int i = 0; i < 3; ++i) axes[i][2] = -axes[i][2];
}

void sampleRigidBodies(const SceneLoader::Scene& scene, const std::string& base_path, std::vector<Vec3>& pos)
{
    for (const SceneLoader::BoundaryData& bd : scene.boundaryModels)
        if (!bd.dynamic) sampleMesh(placedMesh(bd, base_path), scene.particleRadius, pos);
}

std::vector<SceneBody> dynamicRigidBodies(const SceneLoader::Scene& scene, const std::string& base_path)
{
    std::vector<SceneBody> out;
    for (const SceneLoader::BoundaryData& bd : scene.boundaryModels) {
        if (!bd.dynamic) continue;
        const Mesh m = placedMesh(bd, base_path);
        SceneBody b;
        double mass;
        Vec3 com, I;
        Mat3 E;
        meshMassProperties(m.v, m.f, bd.density, mass, com, I, E);
        std::vector<Vec3> world;
        sampleMesh(m, scene.particleRadius, world);
        // body frame: principal axes through the center of mass
        for (const Vec3& p : world) {
            const Vec3 d{p[0] - com[0], p[1] - com[1], p[2] - com[2]};
            b.particles.push_back(Vec3{E[0][0] * d[0] + E[1][0] * d[1] + E[2][0] * d[2],
                                       E[0][1] * d[0] + E[1][1] * d[1] + E[2][1] * d[2],
                                       E[0][2] * d[0] + E[1][2] * d[1] + E[2][2] * d[2]});
        }
        std::memset(&b.body, 0, sizeof b.body);
        b.body.type = DFSPH_BODY_DYNAMIC;
        b.body.mass = mass;
        const std::array<double, 4> q = matQuat(E);
        for (int a = 0; a < 3; ++a) {
            b.body.inertia[a] = I[a];
            b.body.position[a] = com[a];
            b.body.velocity[a] = bd.velocity[a];
            b.body.angular_velocity[a] = bd.angularVelocity[a];
        }
        for (int a = 0; a < 4; ++a) b.body.rotation[a] = q[a];
        out.push_back(std::move(b));
    }
    return out;
}

void sceneParams(const SceneLoader::Scene& scene, dfsph_params& p)
{
    p.particle_radius = scene.particleRadius;
    p.density0 = scene.density0;
    p.viscosity = scene.viscosity;
    for (int a = 0; a < 3; ++a) p.gravity[a] = scene.gravitation[a];
    p.dt = scene.timeStepSize;
    p.max_iter = scene.maxIterations;
    p.max_error = scene.maxError;
    p.max_iter_v = scene.maxIterationsV;
    p.max_error_v = scene.maxErrorV;
    p.use_cfl = scene.cflMethod != 0 ? 1 : 0;
    p.cfl_factor = scene.cflFactor;
    p.dt_max = scene.cflMaxTimeStepSize;
    p.enable_divergence = scene.enableDivergenceSolver ? 1 : 0;
}

// ---------------------------------------------------------------- text files
void writeParticleText(const std::string& path, const std::vector<float>& xyz, const std::vector<float>& vel,
                       const std::vector<float>& mass, int has_factor, int impacted, int dynamic)
{
    const size_t n = mass.size();
    std::ofstream f(path);
    if (!f) throw std::runtime_error("cannot write " + path);
    f << std::setprecision(std::numeric_limits<float>::max_digits10);
    f << n << " " << has_factor << " " << impacted << " " << dynamic << "\n";
    for (size_t i = 0; i < n; ++i) {
        f << mass[i] << " " << xyz[3 * i] << " " << xyz[3 * i + 1] << " " << xyz[3 * i + 2] << " ";
        f << (vel.empty() ? 0.0f : vel[3 * i]) << " " << (vel.empty() ? 0.0f : vel[3 * i + 1]) << " "
          << (vel.empty() ? 0.0f : vel[3 * i + 2]) << " \n";
    }
    if (!f) throw std::runtime_error("write failed: " + path);
}

void readParticleText(const std::string& path, std::vector<float>& xyz, std::vector<float>& vel, std::vector<float>& mass)
{
    std::istringstream in(readFile(path));
    long long n = -1;
    int has_factor = 0, impacted = 0, dynamic = 0;
    in >> n >> has_factor >> impacted >> dynamic;
    if (!in || n < 0) throw std::runtime_error("not a particle file: " + path);
    if (dynamic) {                      // body position, velocity, quaternion, angular velocity
        double skip;
        for (int k = 0; k < 13; ++k) in >> skip;
    }
    xyz.assign(3 * (size_t)n, 0.0f);
    vel.assign(3 * (size_t)n, 0.0f);
    mass.assign((size_t)n, 0.0f);
    for (long long i = 0; i < n; ++i) {
        in >> mass[i] >> xyz[3 * i] >> xyz[3 * i + 1] >> xyz[3 * i + 2] >> vel[3 * i] >> vel[3 * i + 1] >> vel[3 * i + 2];
        if (!in) throw std::runtime_error("truncated particle file: " + path);
    }
}

}  // namespace SPH
