// nifti_io.cpp -- see nifti_io.h.  gzip through zlib's gz* stream API, which
// also reads uncompressed files transparently.
#include "nifti_io.h"

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstring>

namespace bpg_cli {

namespace {

template <typename T>
void swap_bytes(T& v) {
    unsigned char* p = reinterpret_cast<unsigned char*>(&v);
    std::reverse(p, p + sizeof(T));
}

void swap_header(Nifti1Header& h) {
    swap_bytes(h.sizeof_hdr);
    swap_bytes(h.extents);
    swap_bytes(h.session_error);
    for (auto& d : h.dim) swap_bytes(d);
    swap_bytes(h.intent_p1); swap_bytes(h.intent_p2); swap_bytes(h.intent_p3);
    swap_bytes(h.intent_code); swap_bytes(h.datatype); swap_bytes(h.bitpix); swap_bytes(h.slice_start);
    for (auto& d : h.pixdim) swap_bytes(d);
    swap_bytes(h.vox_offset); swap_bytes(h.scl_slope); swap_bytes(h.scl_inter);
    swap_bytes(h.slice_end);
    swap_bytes(h.cal_max); swap_bytes(h.cal_min); swap_bytes(h.slice_duration); swap_bytes(h.toffset);
    swap_bytes(h.glmax); swap_bytes(h.glmin);
    swap_bytes(h.qform_code); swap_bytes(h.sform_code);
    swap_bytes(h.quatern_b); swap_bytes(h.quatern_c); swap_bytes(h.quatern_d);
    swap_bytes(h.qoffset_x); swap_bytes(h.qoffset_y); swap_bytes(h.qoffset_z);
    for (int i = 0; i < 4; i++) { swap_bytes(h.srow_x[i]); swap_bytes(h.srow_y[i]); swap_bytes(h.srow_z[i]); }
}

bool ends_with(const std::string& s, const std::string& t) {
    return s.size() >= t.size() && s.compare(s.size() - t.size(), t.size(), t) == 0;
}

template <typename T>
void convert(const unsigned char* raw, size_t n, bool swap, float* out) {
    for (size_t i = 0; i < n; i++) {
        T v;
        std::memcpy(&v, raw + i * sizeof(T), sizeof(T));
        if (swap) swap_bytes(v);
        out[i] = (float)v;
    }
}

}  // namespace

// the first '.' of the file name (after the last '/')
size_t first_dot(const std::string& name) {
    const size_t slash = name.rfind('/');
    return name.find('.', slash == std::string::npos ? 0 : slash + 1);
}

bool extension_ok(const std::string& name, std::string& ext) {
    const size_t dot = first_dot(name);
    ext = dot == std::string::npos ? std::string() : name.substr(dot);
    return ext == ".nii" || ext == ".nii.gz";
}

std::string insert_suffix(const std::string& name, const std::string& suffix) {
    const size_t dot = first_dot(name);
    if (dot == std::string::npos) return name + suffix;
    return name.substr(0, dot) + suffix + name.substr(dot);
}

bool read_volume(const std::string& path, Volume& out, std::string& err) {
    gzFile f = gzopen(path.c_str(), "rb");
    if (!f) {
        err = "could not open " + path;
        return false;
    }
    Nifti1Header h;
    if (gzread(f, &h, sizeof(h)) != (int)sizeof(h)) {
        gzclose(f);
        err = path + " is too short for a NIfTI-1 header";
        return false;
    }
    bool swap = false;
    if (h.sizeof_hdr != 348) {
        Nifti1Header t = h;
        swap_header(t);
        if (t.sizeof_hdr != 348) {
            gzclose(f);
            err = path + " is not a NIfTI-1 file";
            return false;
        }
        h = t;
        swap = true;
    }
    if (std::memcmp(h.magic, "n+1", 4) != 0) {
        gzclose(f);
        err = path + " is not a single-file NIfTI-1 image (magic n+1)";
        return false;
    }
    const int nd = h.dim[0];
    if (nd < 1 || nd > 7) {
        gzclose(f);
        err = path + ": bad dim[0]";
        return false;
    }
    size_t d[7];
    for (int i = 0; i < 7; i++) d[i] = (i < nd && h.dim[i + 1] > 0) ? (size_t)h.dim[i + 1] : 1;
    out.nx = d[0]; out.ny = d[1]; out.nz = d[2]; out.nt = d[3] * d[4] * d[5] * d[6];
    const size_t n = out.nx * out.ny * out.nz * out.nt;
    size_t bytes = 0;
    switch (h.datatype) {
        case NIFTI_UINT8: bytes = 1; break;
        case NIFTI_INT16: case NIFTI_UINT16: bytes = 2; break;
        case NIFTI_INT32: case NIFTI_FLOAT32: bytes = 4; break;
        case NIFTI_FLOAT64: bytes = 8; break;
        default:
            gzclose(f);
            err = "unknown data type in " + path;
            return false;
    }
    // skip extensions up to vox_offset
    const long off = (long)h.vox_offset;
    if (off > (long)sizeof(h)) {
        std::vector<char> skip((size_t)(off - (long)sizeof(h)));
        if (gzread(f, skip.data(), (unsigned)skip.size()) != (int)skip.size()) {
            gzclose(f);
            err = path + ": truncated before the voxel data";
            return false;
        }
    }
    std::vector<unsigned char> raw(n * bytes);
    size_t got = 0;
    while (got < raw.size()) {
        const unsigned chunk = (unsigned)std::min<size_t>(raw.size() - got, (size_t)1 << 30);
        const int r = gzread(f, raw.data() + got, chunk);
        if (r <= 0) break;
        got += (size_t)r;
    }
    gzclose(f);
    if (got != raw.size()) {
        err = path + ": truncated voxel data";
        return false;
    }
    out.data.resize(n);
    switch (h.datatype) {
        case NIFTI_UINT8: convert<uint8_t>(raw.data(), n, false, out.data.data()); break;
        case NIFTI_INT16: convert<int16_t>(raw.data(), n, swap, out.data.data()); break;
        case NIFTI_UINT16: convert<uint16_t>(raw.data(), n, swap, out.data.data()); break;
        case NIFTI_INT32: convert<int32_t>(raw.data(), n, swap, out.data.data()); break;
        case NIFTI_FLOAT32: convert<float>(raw.data(), n, swap, out.data.data()); break;
        case NIFTI_FLOAT64: convert<double>(raw.data(), n, swap, out.data.data()); break;
    }
    out.hdr = h;
    out.fname = path;
    return true;
}

bool write_volume(const std::string& path, const Volume& like, size_t nt, const float* data, std::string& err) {
    Nifti1Header h = like.hdr;
    const size_t n = like.nx * like.ny * like.nz * nt;
    h.sizeof_hdr = 348;
    h.dim[0] = nt > 1 ? 4 : 3;
    h.dim[1] = (int16_t)like.nx; h.dim[2] = (int16_t)like.ny; h.dim[3] = (int16_t)like.nz;
    h.dim[4] = (int16_t)nt;
    for (int i = 5; i < 8; i++) h.dim[i] = 1;
    h.datatype = NIFTI_FLOAT32;
    h.bitpix = 32;
    h.vox_offset = 352.0f;
    float lo = INFINITY, hi = -INFINITY;
    for (size_t i = 0; i < n; i++) {
        lo = std::min(lo, data[i]);
        hi = std::max(hi, data[i]);
    }
    h.cal_min = n ? lo : 0.0f;
    h.cal_max = n ? hi : 0.0f;
    std::memcpy(h.magic, "n+1", 4);
    const bool gz = ends_with(path, ".gz");
    gzFile f = gzopen(path.c_str(), gz ? "wb6" : "wbT");
    if (!f) {
        err = "could not open " + path + " for writing";
        return false;
    }
    const char ext[4] = {0, 0, 0, 0};  // no extensions
    bool ok = gzwrite(f, &h, sizeof(h)) == (int)sizeof(h) && gzwrite(f, ext, 4) == 4;
    size_t put = 0;
    const unsigned char* src = reinterpret_cast<const unsigned char*>(data);
    while (ok && put < n * sizeof(float)) {
        const unsigned chunk = (unsigned)std::min<size_t>(n * sizeof(float) - put, (size_t)1 << 30);
        ok = gzwrite(f, src + put, chunk) == (int)chunk;
        put += chunk;
    }
    ok = (gzclose(f) == Z_OK) && ok;
    if (!ok) err = "could not write " + path;
    return ok;
}

}  // namespace bpg_cli
