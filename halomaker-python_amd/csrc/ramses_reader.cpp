// ramses_reader.cpp -- RAMSES particle files (part_NNNNN.outCCCCC) read natively, one OpenMP
// thread per file: _read_ramses_new_1010 of src/input_output.py:624-760 (Ra3 and Ra4 layouts,
// families all / dm / star).  Each file is read whole and its Fortran sequential records (a
// 4-byte length before and after each payload) are parsed in place; particles are scattered
// to their dense-id slots as the reference does.
#include <omp.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "ramses_reader.h"

namespace ahop {

namespace {

struct Records {
    std::vector<char> buf;
    size_t pos = 0;
    std::string name;

    void load(const std::string &path)
    {
        name = path;
        FILE *f = std::fopen(path.c_str(), "rb");
        if (!f) throw std::runtime_error("cannot open " + path);
        std::fseek(f, 0, SEEK_END);
        const long n = std::ftell(f);
        std::fseek(f, 0, SEEK_SET);
        buf.resize(n > 0 ? (size_t)n : 0);
        const size_t got = n > 0 ? std::fread(buf.data(), 1, (size_t)n, f) : 0;
        std::fclose(f);
        if (got != buf.size()) throw std::runtime_error("short read of " + path);
        pos = 0;
    }
    // the next record's payload: pointer and byte length (markers checked)
    const char *next(size_t &len)
    {
        if (pos + 4 > buf.size()) throw std::runtime_error(name + ": record past the end");
        uint32_t a, b;
        std::memcpy(&a, buf.data() + pos, 4);
        if (pos + 8 + (size_t)a > buf.size()) throw std::runtime_error(name + ": record past the end");
        std::memcpy(&b, buf.data() + pos + 4 + a, 4);
        if (a != b) throw std::runtime_error(name + ": record markers disagree");
        const char *p = buf.data() + pos + 4;
        pos += 8 + (size_t)a;
        len = a;
        return p;
    }
    void skip(int n = 1)
    {
        size_t len;
        for (int i = 0; i < n; ++i) (void)next(len);
    }
    int32_t one_int()
    {
        size_t len;
        const char *p = next(len);
        if (len < 4) throw std::runtime_error(name + ": short integer record");
        int32_t v;
        std::memcpy(&v, p, 4);
        return v;
    }
    template <typename T>
    const T *array(int64_t n)
    {
        size_t len;
        const char *p = next(len);
        if (len != sizeof(T) * (size_t)n) throw std::runtime_error(name + ": record of unexpected size");
        return reinterpret_cast<const T *>(p);   // payloads are 4-byte aligned at most: read by memcpy
    }
};

template <typename T>
inline T at(const T *p, int64_t i)
{
    T v;
    std::memcpy(&v, reinterpret_cast<const char *>(p) + sizeof(T) * i, sizeof(T));
    return v;
}

std::string part_name(const std::string &repo, const std::string &nchar, int icpu)
{
    char tail[32];
    std::snprintf(tail, sizeof tail, ".out%05d", icpu);
    return repo + "/part_" + nchar + tail;
}

// one file: classify its particles; with outputs, scatter the selected ones
void one_file(const RamsesArgs &A, int icpu, RamsesFileCount &cnt, bool count_only)
{
    Records f;
    f.load(part_name(A.repository, A.nchar, icpu));
    (void)f.one_int();                  // ncpu
    const int ndim = f.one_int();
    const int64_t npart = f.one_int();
    f.skip(1);                          // localseed
    const int nstar = f.one_int();
    f.skip(2);                          // mstar_tot, mstar_lost
    const int nsink = f.one_int();
    if (ndim != 3) throw std::runtime_error(f.name + ": ndim must be 3");
    const double *xp[3] = {nullptr, nullptr, nullptr}, *vp[3] = {nullptr, nullptr, nullptr};
    const double *mp = nullptr;
    if (count_only) {
        f.skip(2 * ndim + 1);
    } else {
        for (int d = 0; d < 3; ++d) xp[d] = f.array<double>(npart);
        for (int d = 0; d < 3; ++d) vp[d] = f.array<double>(npart);
        mp = f.array<double>(npart);
    }
    const int32_t *idp = f.array<int32_t>(npart);
    f.skip(1);                          // level
    const int8_t *fam = nullptr;
    if (A.rver == 4) {
        fam = f.array<int8_t>(npart);
        f.skip(1);                      // tag
    }
    const double *birth = nullptr, *metal = nullptr;
    if (nstar > 0 || nsink > 0) {
        const bool keep = A.stellar && !count_only && A.family != RAMSES_DM;
        if (A.rver == 4) {
            if (keep) {
                birth = f.array<double>(npart);
                metal = f.array<double>(npart);
            }
        } else {
            birth = f.array<double>(npart);   // Ra3 classifies by birth time
            if (keep) metal = f.array<double>(npart);
        }
    }
    auto is_dm = [&](int64_t i) {
        if (A.rver == 4) return at(fam, i) == 1;
        const double t = birth ? at(birth, i) : 0.0;
        return at(idp, i) > 0 && t == 0.0;
    };
    auto is_star = [&](int64_t i) {
        if (A.rver == 4) return at(fam, i) == 2;
        const double t = birth ? at(birth, i) : 0.0;
        const int32_t id = at(idp, i);
        return (t < 0 && id > 0) || (t != 0 && id < 0);
    };
    int64_t ndm = 0, ns = 0, nsel = 0;
    for (int64_t i = 0; i < npart; ++i) {
        const bool d = is_dm(i), s = is_star(i);
        ndm += d;
        ns += s;
        nsel += A.family == RAMSES_ALL ? (d || s) : A.family == RAMSES_DM ? d : s;
    }
    cnt.ndm = ndm;
    cnt.nstar = ns;
    cnt.nselected = nsel;
    if (count_only) return;
    // dense ids (:715-730): stars after the dark matter with family 'all'
    std::vector<int64_t> seen;
    seen.reserve(nsel);
    for (int64_t i = 0; i < npart; ++i) {
        const bool d = is_dm(i), s = is_star(i);
        const bool sel = A.family == RAMSES_ALL ? (d || s) : A.family == RAMSES_DM ? d : s;
        if (!sel) continue;
        const int64_t src = std::llabs((long long)at(idp, i));
        const int64_t local = (A.family == RAMSES_ALL && s) ? src + A.ndm_total : src;
        const int64_t ind = local - 1;
        if (ind < 0 || ind >= A.nselected_total)
            throw std::runtime_error(f.name + ": particle id outside the dense range");
        seen.push_back(ind);
        for (int d3 = 0; d3 < 3; ++d3) {
            A.pos[3 * ind + d3] = at(xp[d3], i) - 0.5;
            A.vel[3 * ind + d3] = at(vp[d3], i) * A.scale_l / A.scale_t * 1e-5;
        }
        A.mass[ind] = at(mp, i);
        if (A.family != RAMSES_ALL && A.source_pid) A.source_pid[ind] = src;
        if (s && A.stellar && A.family != RAMSES_DM && birth && A.birth) {
            const int64_t slot = src - 1;   // star_slots (:746)
            if (slot < 0 || slot >= A.nstar_total)
                throw std::runtime_error(f.name + ": star id outside the star range");
            A.birth[slot] = at(birth, i);
            if (metal && A.metal) A.metal[slot] = at(metal, i);
        }
    }
    // duplicates within the file are an error (:720-723)
    std::vector<int64_t> s2(seen);
    std::sort(s2.begin(), s2.end());
    for (size_t k = 1; k < s2.size(); ++k)
        if (s2[k] == s2[k - 1]) throw std::runtime_error(f.name + ": duplicate particle id");
}

void run_files(const RamsesArgs &A, RamsesFileCount *per_file, bool count_only)
{
    // every file records its own error; the lowest-numbered file's is reported, as the
    // reference, reading the files in order, would stop there
    std::vector<std::string> err(A.ncpu);
    const int nt = A.nthreads > 0 ? A.nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int icpu = 1; icpu <= A.ncpu; ++icpu) {
        try {
            one_file(A, icpu, per_file[icpu - 1], count_only);
        } catch (const std::exception &ex) {
            err[icpu - 1] = ex.what();
        }
    }
    for (const std::string &e : err)
        if (!e.empty()) throw std::runtime_error(e);
}

}  // namespace

void ramses_count(const RamsesArgs &A, RamsesFileCount *per_file)
{
    if (A.ncpu < 1) throw std::runtime_error("ncpu must be >= 1");
    if (A.rver != 3 && A.rver != 4) throw std::runtime_error("rver must be 3 (Ra3) or 4 (Ra4)");
    run_files(A, per_file, true);
}

void ramses_read(const RamsesArgs &A, RamsesFileCount *per_file)
{
    if (A.ncpu < 1) throw std::runtime_error("ncpu must be >= 1");
    if (A.rver != 3 && A.rver != 4) throw std::runtime_error("rver must be 3 (Ra3) or 4 (Ra4)");
    if (!A.pos || !A.vel || !A.mass) throw std::runtime_error("null output array");
    run_files(A, per_file, false);
}

}  // namespace ahop
