// Restart files of the dynamics state in the layout of the reference's restart writer,
// ParaGridIO::dumpModelState (core/src/ParaGridIO.cpp:272-349), as a netCDF classic file
// with 64-bit offsets (CDF-2), written and read here without the netCDF library.
//
// The reference writes netCDF-4 with groups: a "structure" group and a "metadata" group
// whose "type" attribute is the structure name (CommonRestartMetadata.cpp:18-31), and a
// "data" group with every defined dimension (ModelArrayDetails.cpp:28-41) and one FP64
// variable per field with a "missing_value" attribute. CDF-2 has no groups, so the data
// group's dimensions and variables sit at the root and the metadata become global
// attributes. The dimension names, their order, the variables' dimension order (reversed
// ModelArray dimensions, components fastest) and the memory layout are the reference's:
//   H        (ydim, xdim)                  element ix + nx iy
//   DG       (ydim, xdim, dg_comp)         ModelArray rows, components contiguous
//   DGSTRESS (ydim, xdim, dgstress_comp)
//   CG       (y_cg, x_cg)                  node ix + (CG nx + 1) iy
//   VERTEX   (yvertex, xvertex, ncoords)
// Fields: hice, cice (DG); u, v (H, the cell means the reference restarts from); mask (H);
// coords (VERTEX); and, to restore this library's device state exactly, the stress and
// strain s11 ... e22 (DGSTRESS) and the CG velocities u_cg, v_cg; the brittle kernel adds
// damage (DG) and its step-averaged velocity avgU_cg, avgV_cg (CG).
#include "nsdg_dyn.hpp"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <map>

namespace nsdg {
namespace {

constexpr uint32_t kDimTag = 0x0A, kVarTag = 0x0B, kAttTag = 0x0C;
constexpr uint32_t kChar = 2, kInt = 4, kDouble = 6;
constexpr double kMissing = 1.7e38; // MissingData::defaultValue (core/src/include/MissingData.hpp:17)

struct Writer {
    std::vector<uint8_t> b;
    void u32(uint32_t v)
    {
        for (int k = 3; k >= 0; --k)
            b.push_back(static_cast<uint8_t>(v >> (8 * k)));
    }
    void u64(uint64_t v)
    {
        for (int k = 7; k >= 0; --k)
            b.push_back(static_cast<uint8_t>(v >> (8 * k)));
    }
    void f64(double d)
    {
        uint64_t v;
        std::memcpy(&v, &d, 8);
        u64(v);
    }
    void pad()
    {
        while (b.size() % 4)
            b.push_back(0);
    }
    void name(const std::string& s)
    {
        u32(static_cast<uint32_t>(s.size()));
        b.insert(b.end(), s.begin(), s.end());
        pad();
    }
};

struct Attr {
    std::string name;
    uint32_t type;
    std::string text;       // kChar
    std::vector<double> d;  // kDouble
    std::vector<int32_t> i; // kInt
};
void putAttrs(Writer& w, const std::vector<Attr>& at)
{
    if (at.empty()) {
        w.u32(0);
        w.u32(0);
        return;
    }
    w.u32(kAttTag);
    w.u32(static_cast<uint32_t>(at.size()));
    for (const auto& a : at) {
        w.name(a.name);
        w.u32(a.type);
        if (a.type == kChar) {
            w.u32(static_cast<uint32_t>(a.text.size()));
            w.b.insert(w.b.end(), a.text.begin(), a.text.end());
            w.pad();
        } else if (a.type == kDouble) {
            w.u32(static_cast<uint32_t>(a.d.size()));
            for (double v : a.d)
                w.f64(v);
        } else {
            w.u32(static_cast<uint32_t>(a.i.size()));
            for (int32_t v : a.i)
                w.u32(static_cast<uint32_t>(v));
        }
    }
}

struct Dim {
    std::string name;
    size_t len;
};
struct Var {
    std::string name;
    std::vector<int> dims;
    std::vector<double> data;
};

// Dimension ids in the order of ModelArray::definedDimensions (ModelArrayDetails.cpp:28-41).
enum { D_X, D_Y, D_Z, D_XV, D_YV, D_XCG, D_YCG, D_DG, D_DGS, D_NC, D_COUNT };

struct Layout {
    size_t nx, ny, ncg_x, ncg_y;
    int cg, dg, s;
    bool brittle;
    std::vector<Dim> dims() const
    {
        return { { "xdim", nx }, { "ydim", ny }, { "zdim", 1 }, { "xvertex", nx + 1 }, { "yvertex", ny + 1 },
            { "x_cg", ncg_x }, { "y_cg", ncg_y }, { "dg_comp", static_cast<size_t>(dg) },
            { "dgstress_comp", static_cast<size_t>(s) }, { "ncoords", 2 } };
    }
};
Layout layoutOf(DynBase& k)
{
    const HostMesh& m = k.meshHost();
    Layout L;
    L.nx = m.nx;
    L.ny = m.ny;
    L.cg = k.cgDegree();
    L.dg = k.dgComp();
    L.s = k.stressComp();
    L.brittle = k.brittleRheology();
    L.ncg_x = L.cg * L.nx + 1;
    L.ncg_y = L.cg * L.ny + 1;
    return L;
}

const std::vector<int> kH { D_Y, D_X }, kDG { D_Y, D_X, D_DG }, kDGS { D_Y, D_X, D_DGS }, kCG { D_YCG, D_XCG },
    kVertex { D_YV, D_XV, D_NC };

// ---------------------------------------------------------------- reading
struct Reader {
    const std::vector<uint8_t>& b;
    size_t p = 0;
    int version = 0;
    void need(size_t n) const { require(p + n <= b.size(), "restart file truncated"); }
    uint32_t u32()
    {
        need(4);
        uint32_t v = 0;
        for (int k = 0; k < 4; ++k)
            v = (v << 8) | b[p++];
        return v;
    }
    uint64_t u64()
    {
        need(8);
        uint64_t v = 0;
        for (int k = 0; k < 8; ++k)
            v = (v << 8) | b[p++];
        return v;
    }
    std::string name()
    {
        const uint32_t n = u32();
        need(n);
        std::string s(reinterpret_cast<const char*>(&b[p]), n);
        p += (n + 3) / 4 * 4;
        return s;
    }
    static size_t typeSize(uint32_t t)
    {
        switch (t) {
        case 1:
        case 2: return 1;
        case 3: return 2;
        case 4:
        case 5: return 4;
        case 6: return 8;
        default: throw Error(NSDG_ERR_ARG, "restart file: unknown netCDF type");
        }
    }
    void skipAttrs(std::vector<std::pair<std::string, std::string>>* text = nullptr)
    {
        const uint32_t tag = u32(), n = u32();
        require(tag == kAttTag || (tag == 0 && n == 0), "restart file: bad attribute list");
        for (uint32_t k = 0; k < n; ++k) {
            const std::string nm = name();
            const uint32_t t = u32(), cnt = u32();
            const size_t bytes = (typeSize(t) * cnt + 3) / 4 * 4;
            need(bytes);
            if (text && t == kChar)
                text->push_back({ nm, std::string(reinterpret_cast<const char*>(&b[p]), cnt) });
            p += bytes;
        }
    }
};

} // namespace

void writeRestart(DynBase& k, const char* path, double timeSeconds)
{
    require(path != nullptr, "path is NULL");
    const Layout L = layoutOf(k);
    const size_t ne = L.nx * L.ny, ncg = L.ncg_x * L.ncg_y;
    std::vector<Var> vars;
    auto dgField = [&](const char* n, int nc, const std::vector<int>& dims) {
        Var v { n, dims, std::vector<double>(ne * nc) };
        k.getField(n, v.data.data(), nc);
        vars.push_back(std::move(v));
    };
    auto cgField = [&](const char* src, const char* n) {
        Var v { n, kCG, std::vector<double>(ncg) };
        k.getCG(src, v.data.data());
        vars.push_back(std::move(v));
    };
    dgField("hice", L.dg, kDG);
    dgField("cice", L.dg, kDG);
    dgField("u", 1, kH);
    dgField("v", 1, kH);
    if (L.brittle)
        dgField("damage", L.dg, kDG);
    for (const char* n : { "s11", "s12", "s22", "e11", "e12", "e22" })
        dgField(n, L.s, kDGS);
    cgField("u", "u_cg");
    cgField("v", "v_cg");
    if (L.brittle) {
        cgField("avgU", "avgU_cg");
        cgField("avgV", "avgV_cg");
    }
    const HostMesh& m = k.meshHost();
    {
        Var v { "mask", kH, std::vector<double>(ne) };
        for (size_t e = 0; e < ne; ++e)
            v.data[e] = m.ocean[e] ? 1.0 : 0.0;
        vars.push_back(std::move(v));
        vars.push_back({ "coords", kVertex, m.coords });
    }

    const std::vector<Dim> dims = L.dims();
    std::vector<Attr> gatts {
        { "type", kChar, "parametric_rectangular", {}, {} },
        { "time", kDouble, "", { timeSeconds }, {} },
        { "time_units", kChar, "seconds since the start of the run", {}, {} },
        { "rheology", kChar, L.brittle ? "bbm" : "mevp", {}, {} },
        { "cg_degree", kInt, "", {}, { L.cg } },
        { "coordinate_system", kChar,
            m.spherical ? "spherical, radians, after the pole rotation of the mesh" : "cartesian, metres", {}, {} },
        { "layout", kChar,
            "ParaGridIO::dumpModelState data group (netCDF classic: groups flattened, metadata as attributes)",
            {}, {} },
    };
    const std::vector<Attr> vatt { { "missing_value", kDouble, "", { kMissing }, {} } };

    // header with placeholder offsets, then the data of each variable
    Writer w;
    w.b = { 'C', 'D', 'F', 2 };
    w.u32(0); // numrecs
    w.u32(kDimTag);
    w.u32(static_cast<uint32_t>(dims.size()));
    for (const auto& d : dims) {
        w.name(d.name);
        w.u32(static_cast<uint32_t>(d.len));
    }
    putAttrs(w, gatts);
    w.u32(kVarTag);
    w.u32(static_cast<uint32_t>(vars.size()));
    std::vector<size_t> beginAt(vars.size());
    for (size_t i = 0; i < vars.size(); ++i) {
        const Var& v = vars[i];
        size_t count = 1;
        for (int d : v.dims)
            count *= dims[d].len;
        require(count == v.data.size(), "restart: field size does not match its dimensions");
        w.name(v.name);
        w.u32(static_cast<uint32_t>(v.dims.size()));
        for (int d : v.dims)
            w.u32(static_cast<uint32_t>(d));
        putAttrs(w, vatt);
        w.u32(kDouble);
        const uint64_t vsize = count * 8;
        // CDF-2 sizes each variable in 32 bits (only a last variable may exceed it)
        require(vsize < 0xFFFFFFFCull, "restart: field " + v.name + " exceeds the 4 GiB a netCDF classic variable holds",
            NSDG_ERR_UNSUPPORTED);
        w.u32(static_cast<uint32_t>(vsize));
        beginAt[i] = w.b.size();
        w.u64(0);
    }
    for (size_t i = 0; i < vars.size(); ++i) {
        const uint64_t begin = w.b.size();
        for (int k8 = 0; k8 < 8; ++k8)
            w.b[beginAt[i] + k8] = static_cast<uint8_t>(begin >> (8 * (7 - k8)));
        for (double d : vars[i].data)
            w.f64(d);
    }
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    require(static_cast<bool>(f), std::string("cannot open ") + path + " for writing");
    f.write(reinterpret_cast<const char*>(w.b.data()), static_cast<std::streamsize>(w.b.size()));
    require(static_cast<bool>(f), std::string("write failed: ") + path);
}

void readRestart(DynBase& k, const char* path, double* timeSeconds)
{
    require(path != nullptr, "path is NULL");
    std::ifstream f(path, std::ios::binary);
    require(static_cast<bool>(f), std::string("cannot open ") + path);
    const std::vector<uint8_t> bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    Reader r { bytes };
    r.need(4);
    require(bytes[0] == 'C' && bytes[1] == 'D' && bytes[2] == 'F' && (bytes[3] == 1 || bytes[3] == 2),
        "not a netCDF classic file");
    r.version = bytes[3];
    r.p = 4;
    require(r.u32() == 0, "restart file: record variables are not used here");
    std::vector<Dim> dims;
    {
        const uint32_t tag = r.u32(), n = r.u32();
        require(tag == kDimTag || (tag == 0 && n == 0), "restart file: bad dimension list");
        for (uint32_t i = 0; i < n; ++i) {
            std::string nm = r.name();
            dims.push_back({ nm, r.u32() });
        }
    }
    // global attributes: the time and the rheology
    double t = 0;
    std::string rheology;
    {
        const uint32_t tag = r.u32(), n = r.u32();
        require(tag == kAttTag || (tag == 0 && n == 0), "restart file: bad attribute list");
        for (uint32_t i = 0; i < n; ++i) {
            const std::string nm = r.name();
            const uint32_t ty = r.u32(), cnt = r.u32();
            const size_t nb = (Reader::typeSize(ty) * cnt + 3) / 4 * 4;
            r.need(nb);
            if (nm == "time" && cnt == 1 && ty >= 3 && ty <= 6) { // any numeric type
                Reader sub { bytes };
                sub.p = r.p;
                if (ty == kDouble) {
                    const uint64_t v = sub.u64();
                    std::memcpy(&t, &v, 8);
                } else if (ty == 5) {
                    const uint32_t v = sub.u32();
                    float fv;
                    std::memcpy(&fv, &v, 4);
                    t = fv;
                } else if (ty == kInt)
                    t = static_cast<int32_t>(sub.u32());
                else
                    t = static_cast<int16_t>((bytes[sub.p] << 8) | bytes[sub.p + 1]);
            } else if (nm == "rheology" && ty == kChar)
                rheology.assign(reinterpret_cast<const char*>(&bytes[r.p]), cnt);
            r.p += nb;
        }
    }
    struct VarInfo {
        std::vector<uint32_t> dims;
        uint32_t type;
        uint64_t begin;
    };
    std::map<std::string, VarInfo> vars;
    {
        const uint32_t tag = r.u32(), n = r.u32();
        require(tag == kVarTag || (tag == 0 && n == 0), "restart file: bad variable list");
        for (uint32_t i = 0; i < n; ++i) {
            VarInfo vi;
            const std::string nm = r.name();
            const uint32_t nd = r.u32();
            for (uint32_t j = 0; j < nd; ++j) {
                vi.dims.push_back(r.u32());
                require(vi.dims.back() < dims.size(), "restart file: bad dimension id");
            }
            r.skipAttrs();
            vi.type = r.u32();
            (void)r.u32(); // vsize
            vi.begin = r.version == 2 ? r.u64() : r.u32();
            vars[nm] = vi;
        }
    }
    const Layout L = layoutOf(k);
    const std::vector<Dim> want = L.dims();
    auto dimLen = [&](const std::string& nm) -> size_t {
        for (const auto& d : dims)
            if (d.name == nm)
                return d.len;
        throw Error(NSDG_ERR_ARG, "restart file has no dimension " + nm);
    };
    for (const char* nm : { "xdim", "ydim", "x_cg", "y_cg", "dg_comp", "dgstress_comp" }) {
        size_t w = 0;
        for (const auto& d : want)
            if (d.name == nm)
                w = d.len;
        require(dimLen(nm) == w,
            std::string("restart file dimension ") + nm + " does not match this kernel (" + std::to_string(dimLen(nm))
                + " against " + std::to_string(w) + ")");
    }
    require(rheology == (L.brittle ? "bbm" : "mevp"), "restart file of the other rheology (" + rheology + ")");
    auto load = [&](const std::string& nm, const std::vector<std::string>& dimNames) {
        auto it = vars.find(nm);
        require(it != vars.end(), "restart file has no variable " + nm);
        const VarInfo& vi = it->second;
        require(vi.type == kDouble, "restart variable " + nm + " is not double");
        require(vi.dims.size() == dimNames.size(), "restart variable " + nm + " has the wrong rank");
        size_t count = 1;
        for (size_t j = 0; j < dimNames.size(); ++j) {
            require(dims[vi.dims[j]].name == dimNames[j], "restart variable " + nm + " has the wrong dimensions");
            count *= dims[vi.dims[j]].len;
        }
        require(vi.begin + count * 8 <= bytes.size(), "restart variable " + nm + " truncated");
        std::vector<double> out(count);
        Reader sub { bytes };
        sub.p = vi.begin;
        for (size_t i = 0; i < count; ++i) {
            const uint64_t v = sub.u64();
            std::memcpy(&out[i], &v, 8);
        }
        return out;
    };
    const std::vector<std::string> dg { "ydim", "xdim", "dg_comp" }, dgs { "ydim", "xdim", "dgstress_comp" },
        cg { "y_cg", "x_cg" };
    // read everything first, then set the device state
    const auto h = load("hice", dg), a = load("cice", dg);
    std::vector<std::vector<double>> tensors;
    for (const char* n : { "s11", "s12", "s22", "e11", "e12", "e22" })
        tensors.push_back(load(n, dgs));
    const auto ucg = load("u_cg", cg), vcg = load("v_cg", cg);
    std::vector<double> dmg, au, av;
    if (L.brittle) {
        dmg = load("damage", dg);
        au = load("avgU_cg", cg);
        av = load("avgV_cg", cg);
    }
    k.setField("hice", h.data(), L.dg);
    k.setField("cice", a.data(), L.dg);
    const char* tn[] = { "s11", "s12", "s22", "e11", "e12", "e22" };
    for (int i = 0; i < 6; ++i)
        k.setField(tn[i], tensors[i].data(), L.s);
    k.setCG("u", ucg.data());
    k.setCG("v", vcg.data());
    if (L.brittle) {
        k.setField("damage", dmg.data(), L.dg);
        k.setCG("avgU", au.data());
        k.setCG("avgV", av.data());
    }
    if (timeSeconds)
        *timeSeconds = t;
}

} // namespace nsdg
