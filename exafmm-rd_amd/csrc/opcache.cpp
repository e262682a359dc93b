// On-disk cache of the precomputed operators of one (P, images, cycle): the reference recomputes
// its SVD, M2M/L2L and M2L kernels on every solve (3dpp/src/kernel.cpp:1638-1850); a plan can
// instead load them from a file written by an earlier plan with the same configuration.
//
// File layout (little endian): magic "FMM3DPP-OPS\0", u32 version, i32 P, i32 images, f64 cycle,
// u32 array count, then per array: u32 name length, name bytes, u8 type (0 f64, 1 i32), u64
// element count, elements; finally a u64 FNV-1a hash of every byte before it.  Every field of
// Operators is stored, the derived device forms included, so a loaded set is bit-identical to the
// one that was saved.
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <type_traits>
#include <string>
#include <vector>

#include "structure.h"

namespace fmm3dpp {

namespace {

constexpr char kMagic[12] = "FMM3DPP-OPS";
// Raised whenever build_operators' results change, so that files of an older computation (in an
// FMM3DPP_OPERATOR_CACHE directory, say) are refused rather than used.
constexpr uint32_t kVersion = 1;

struct Hash {
    uint64_t h = 1469598103934665603ULL;
    void add(const void* p, size_t n) {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < n; i++) {
            h ^= c[i];
            h *= 1099511628211ULL;
        }
    }
};

struct Writer {
    FILE* f;
    Hash hash;
    bool ok = true;
    void put(const void* p, size_t n) {
        if (!ok) return;
        ok = fwrite(p, 1, n, f) == n;
        hash.add(p, n);
    }
    template <class T> void val(T v) { put(&v, sizeof(T)); }
    template <class T> void arr(const std::string& name, const std::vector<T>& a) {
        val<uint32_t>((uint32_t)name.size());
        put(name.data(), name.size());
        val<uint8_t>(std::is_same<T, double>::value ? 0 : 1);
        val<uint64_t>(a.size());
        put(a.data(), a.size() * sizeof(T));
    }
};

struct Reader {
    FILE* f;
    Hash hash;
    bool ok = true;
    void get(void* p, size_t n) {
        if (!ok) return;
        ok = fread(p, 1, n, f) == n;
        if (ok) hash.add(p, n);
    }
    template <class T> T val() {
        T v{};
        get(&v, sizeof(T));
        return v;
    }
    // `expect`: the element count the plan's configuration implies; any other count in the file is
    // refused before anything is allocated for it.
    template <class T> bool arr(const std::string& name, std::vector<T>& a, uint64_t expect) {
        const uint32_t len = val<uint32_t>();
        if (!ok || len != name.size()) return ok = false;
        std::string got(len, '\0');
        get(&got[0], len);
        const uint8_t type = val<uint8_t>();
        const uint64_t n = val<uint64_t>();
        if (!ok || got != name || type != (std::is_same<T, double>::value ? 0 : 1) || n != expect)
            return ok = false;
        a.resize(n);
        get(a.data(), n * sizeof(T));
        return ok;
    }
};

// The arrays in file order.
template <class F>
void each_array(Operators& o, F&& fn) {
    fn("UT", o.UT);
    fn("VSinv", o.VSinv);
    fn("UT_l2p", o.UT_l2p);
    fn("VSinv_l2p", o.VSinv_l2p);
    for (int k = 0; k < 8; k++) fn("M2M" + std::to_string(k), o.m2m[k]);
    for (int k = 0; k < 27; k++) fn("M2M_IMG" + std::to_string(k), o.m2m_img[k]);
    for (int k = 0; k < 8; k++) fn("L2L" + std::to_string(k), o.l2l[k]);
    fn("L2L_IMG", o.l2l_img);
    fn("GK8", o.gk8);
    fn("GK27", o.gk27);
    fn("M2M_CAT", o.m2m_cat);
    fn("M2M_IMG_SUM", o.m2m_img_sum);
    fn("GK27_SUM", o.gk27_sum);
    fn("UNIT_SURF", o.unit_surf);
    fn("LATTICE", o.lattice);
    fn("TWIDDLE", o.twiddle);
}
constexpr uint32_t kArrays = 4 + 8 + 27 + 8 + 1 + 2 + 3 + 3;

// Element count of each array for (P, images), in the layouts of build_operators.
uint64_t expected_size(const std::string& name, int P, int images) {
    const uint64_t S = (uint64_t)surface_count(P), N = 2 * (uint64_t)P - 1, F = N * N * (N / 2 + 1);
    const uint64_t SS = S * S, SSi = images > 0 ? SS : 0;
    if (name == "GK8") return 26 * F * 128;
    if (name == "GK27") return images >= 2 ? 26 * F * 56 : 0;
    if (name == "GK27_SUM") return 26 * F * 2;
    if (name == "M2M_CAT") return 8 * SS;
    if (name == "M2M_IMG_SUM") return SS;
    if (name == "UNIT_SURF" || name == "LATTICE") return 3 * S;
    if (name == "TWIDDLE") return 2 * N;
    if (name == "L2L_IMG" || name.rfind("M2M_IMG", 0) == 0) return SSi;
    return SS;   // UT, VSinv, UT_l2p, VSinv_l2p, M2M0..7, L2L0..7
}

}  // namespace

int save_operators(const Operators& ops, double cycle, const std::string& path, std::string& err) {
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) { err = "cannot open " + path + " for writing"; return FMM3DPP_ERR_ARG; }
    Writer w{f, Hash{}};
    w.put(kMagic, sizeof(kMagic));
    w.val<uint32_t>(kVersion);
    w.val<int32_t>(ops.P);
    w.val<int32_t>(ops.images);
    w.val<double>(cycle);
    w.val<uint32_t>(kArrays);
    each_array(const_cast<Operators&>(ops), [&](const std::string& name, auto& a) { w.arr(name, a); });
    const uint64_t h = w.hash.h;
    const bool ok = w.ok && fwrite(&h, 1, 8, f) == 8;
    if (fclose(f) != 0 || !ok) { err = "write to " + path + " failed"; return FMM3DPP_ERR_ARG; }
    return FMM3DPP_OK;
}

int load_operators(Operators& ops, int P, int images, double cycle, const std::string& path, std::string& err) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) { err = "cannot open " + path; return FMM3DPP_ERR_ARG; }
    Reader r{f, Hash{}};
    char magic[sizeof(kMagic)];
    r.get(magic, sizeof(magic));
    const uint32_t version = r.val<uint32_t>();
    const int32_t fP = r.val<int32_t>(), fimages = r.val<int32_t>();
    const double fcycle = r.val<double>();
    const uint32_t narr = r.val<uint32_t>();
    auto fail = [&](const std::string& m) { fclose(f); err = path + ": " + m; return FMM3DPP_ERR_ARG; };
    if (!r.ok || memcmp(magic, kMagic, sizeof(kMagic)) != 0) return fail("not an operator file");
    if (version != kVersion || narr != kArrays) return fail("unsupported operator file version");
    if (fP != P || fimages != images || fcycle != cycle)
        return fail("operators are for P=" + std::to_string(fP) + " images=" + std::to_string(fimages) +
                    ", not for this plan's configuration");
    Operators o;
    try {
        each_array(o, [&](const std::string& name, auto& a) { r.arr(name, a, expected_size(name, P, images)); });
    } catch (const std::bad_alloc&) {
        return fail("not enough host memory for the operators");
    }
    const uint64_t want = r.hash.h;
    uint64_t got = 0;
    const bool tail = r.ok && fread(&got, 1, 8, f) == 8;
    fclose(f);
    if (!tail || got != want) { err = path + ": truncated or corrupt operator file"; return FMM3DPP_ERR_ARG; }
    o.P = P;
    o.S = surface_count(P);
    o.N = 2 * P - 1;
    o.Nh = o.N / 2 + 1;
    o.F = o.N * o.N * o.Nh;
    o.images = images;
    o.r0 = cycle / 2;
    const size_t SS = (size_t)o.S * o.S, SSi = images > 0 ? SS : 0;   // image operators exist for images >= 1
    bool sizes = o.UT.size() == SS && o.VSinv.size() == SS && o.UT_l2p.size() == SS && o.VSinv_l2p.size() == SS &&
                 o.l2l_img.size() == SSi && o.gk8.size() == (size_t)26 * o.F * 128 &&
                 o.gk27.size() == (images >= 2 ? (size_t)26 * o.F * 56 : 0) &&
                 o.gk27_sum.size() == (size_t)26 * o.F * 2 &&
                 o.m2m_img_sum.size() == SS && o.m2m_cat.size() == 8 * SS &&
                 o.unit_surf.size() == (size_t)3 * o.S && o.lattice.size() == (size_t)3 * o.S &&
                 o.twiddle.size() == (size_t)2 * o.N;
    for (int k = 0; k < 8; k++) sizes = sizes && o.m2m[k].size() == SS && o.l2l[k].size() == SS;
    for (int k = 0; k < 27; k++) sizes = sizes && o.m2m_img[k].size() == SSi;
    if (!sizes) { err = path + ": operator sizes do not match P=" + std::to_string(P); return FMM3DPP_ERR_ARG; }
    ops = std::move(o);
    return FMM3DPP_OK;
}

int build_operators_cached(int P, int images, double cycle, Operators& ops, std::string& err) {
    const char* dir = std::getenv("FMM3DPP_OPERATOR_CACHE");
    if (!dir || !*dir) return build_operators(P, images, cycle, ops, err);
    uint64_t cbits;
    std::memcpy(&cbits, &cycle, 8);
    char name[128];
    snprintf(name, sizeof(name), "/fmm3dpp_ops_v%u_P%d_i%d_c%016llx.bin", (unsigned)kVersion, P, images,
             (unsigned long long)cbits);
    const std::string path = std::string(dir) + name;
    std::string ignored;
    if (load_operators(ops, P, images, cycle, path, ignored) == FMM3DPP_OK) return FMM3DPP_OK;
    const int rc = build_operators(P, images, cycle, ops, err);
    if (rc != FMM3DPP_OK) return rc;
    // written under a name of this process, then renamed: a reader sees no file or a whole one
    const std::string tmp = path + ".tmp" + std::to_string((long long)getpid());
    if (save_operators(ops, cycle, tmp, ignored) != FMM3DPP_OK || std::rename(tmp.c_str(), path.c_str()) != 0)
        std::remove(tmp.c_str());
    return FMM3DPP_OK;
}

}  // namespace fmm3dpp
