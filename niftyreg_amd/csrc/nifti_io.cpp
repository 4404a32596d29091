// The file formats at the edges of a registration: single-file NIfTI-1 images (.nii, .nii.gz)
// and the affine text matrix.  Host code only: the registrations take and return plain arrays,
// and these entry points turn files into those arrays and back the way the reference's
// reg_io_ReadImageFile / reg_io_WriteImageFile (reg-io/_reg_ReadWriteImage.cpp:51-81,145-170,
// over niftilib) and reg_tool_{Read,Write}AffineFile (reg-io/_reg_ReadWriteMatrix.cpp:83-116) do.
//
// The header is the 348-byte NIfTI-1 layout (a published format, so the field offsets are the
// standard's); geometry follows niftilib's rules: the qform from the quaternion (or the pixel
// spacings when qform_code <= 0), the sform from its rows, inverses by the adjugate with a
// long-double reciprocal, then reg_checkAndCorrectDimension (reg-lib/cpu/_reg_tools.cpp:17-71).

#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "host.hpp"

namespace nrh {
namespace {

struct mat44f {
    float m[4][4];
};

#pragma pack(push, 1)
struct Nifti1Header {
    int32_t sizeof_hdr;
    char data_type[10];
    char db_name[18];
    int32_t extents;
    int16_t session_error;
    char regular;
    char dim_info;
    int16_t dim[8];
    float intent_p1, intent_p2, intent_p3;
    int16_t intent_code;
    int16_t datatype;
    int16_t bitpix;
    int16_t slice_start;
    float pixdim[8];
    float vox_offset;
    float scl_slope, scl_inter;
    int16_t slice_end;
    char slice_code;
    char xyzt_units;
    float cal_max, cal_min;
    float slice_duration, toffset;
    int32_t glmax, glmin;
    char descrip[80];
    char aux_file[24];
    int16_t qform_code, sform_code;
    float quatern_b, quatern_c, quatern_d;
    float qoffset_x, qoffset_y, qoffset_z;
    float srow_x[4], srow_y[4], srow_z[4];
    char intent_name[16];
    char magic[4];
};
#pragma pack(pop)
static_assert(sizeof(Nifti1Header) == 348, "NIfTI-1 header is 348 bytes");

template<class T> void swap_bytes(T& v) {
    unsigned char *p = reinterpret_cast<unsigned char *>(&v);
    for (size_t i = 0; i < sizeof(T) / 2; ++i) std::swap(p[i], p[sizeof(T) - 1 - i]);
}

void swap_header(Nifti1Header& h) {
    swap_bytes(h.sizeof_hdr);
    swap_bytes(h.extents);
    swap_bytes(h.session_error);
    for (auto& d : h.dim) swap_bytes(d);
    swap_bytes(h.intent_p1);
    swap_bytes(h.intent_p2);
    swap_bytes(h.intent_p3);
    swap_bytes(h.intent_code);
    swap_bytes(h.datatype);
    swap_bytes(h.bitpix);
    swap_bytes(h.slice_start);
    for (auto& p : h.pixdim) swap_bytes(p);
    swap_bytes(h.vox_offset);
    swap_bytes(h.scl_slope);
    swap_bytes(h.scl_inter);
    swap_bytes(h.slice_end);
    swap_bytes(h.cal_max);
    swap_bytes(h.cal_min);
    swap_bytes(h.slice_duration);
    swap_bytes(h.toffset);
    swap_bytes(h.glmax);
    swap_bytes(h.glmin);
    swap_bytes(h.qform_code);
    swap_bytes(h.sform_code);
    swap_bytes(h.quatern_b);
    swap_bytes(h.quatern_c);
    swap_bytes(h.quatern_d);
    swap_bytes(h.qoffset_x);
    swap_bytes(h.qoffset_y);
    swap_bytes(h.qoffset_z);
    for (int i = 0; i < 4; ++i) {
        swap_bytes(h.srow_x[i]);
        swap_bytes(h.srow_y[i]);
        swap_bytes(h.srow_z[i]);
    }
}

int bytes_of(int datatype) {
    switch (datatype) {
    case 2: case 256: return 1;             // uint8, int8
    case 4: case 512: return 2;             // int16, uint16
    case 8: case 16: case 768: return 4;    // int32, float32, uint32
    case 64: case 1024: case 1280: return 8;  // float64, int64, uint64
    default: return 0;
    }
}

struct GzFile {
    gzFile f = nullptr;
    ~GzFile() {
        if (f) gzclose(f);
    }
};

// The header of a single-file NIfTI-1 image, byte order corrected; swapped = file was foreign-endian.
Nifti1Header read_raw_header(const char *path, GzFile& gz, bool& swapped) {
    gz.f = gzopen(path, "rb");
    NRH_REQUIRE(gz.f, NRH_ERR_ARG, std::string("cannot open ") + path);
    Nifti1Header h;
    NRH_REQUIRE(gzread(gz.f, &h, sizeof(h)) == (int)sizeof(h), NRH_ERR_ARG, std::string("short header in ") + path);
    swapped = false;
    if (h.sizeof_hdr != 348) {
        swap_header(h);
        swapped = true;
    }
    NRH_REQUIRE(h.sizeof_hdr == 348, NRH_ERR_ARG, std::string("not a NIfTI-1 file: ") + path);
    NRH_REQUIRE(std::memcmp(h.magic, "n+1", 4) == 0, NRH_ERR_UNSUPPORTED,
                std::string("only single-file NIfTI-1 images (.nii, .nii.gz) are read: ") + path);
    NRH_REQUIRE(h.dim[0] >= 1 && h.dim[0] <= 7, NRH_ERR_ARG, "bad dim[0]");
    NRH_REQUIRE(bytes_of(h.datatype) > 0, NRH_ERR_UNSUPPORTED, "unsupported NIfTI datatype");
    return h;
}

mat44f quatern_to_mat44(float qb, float qc, float qd, float qx, float qy, float qz, float dx, float dy, float dz,
                        float qfac) {
    // the NIfTI-1 quaternion parametrisation (a = sqrt(1 - b^2 - c^2 - d^2), 180-degree case when
    // that is below 1e-7), with niftilib's double / long-double arithmetic
    mat44f R{};
    double a, b = qb, c = qc, d = qd, xd, yd, zd;
    a = (double)(1.0L - (long double)(b * b + c * c + d * d));
    if (a < 1.e-7) {
        a = (double)(1.0L / (long double)std::sqrt(b * b + c * c + d * d));
        b *= a;
        c *= a;
        d *= a;
        a = 0.0;
    } else {
        a = std::sqrt(a);
    }
    xd = dx > 0.0f ? dx : 1.0;
    yd = dy > 0.0f ? dy : 1.0;
    zd = dz > 0.0f ? dz : 1.0;
    if (qfac < 0.0f) zd = -zd;
    R.m[0][0] = (float)((a * a + b * b - c * c - d * d) * xd);
    R.m[0][1] = (float)(2.0L * (long double)(b * c - a * d) * (long double)yd);
    R.m[0][2] = (float)(2.0L * (long double)(b * d + a * c) * (long double)zd);
    R.m[1][0] = (float)(2.0L * (long double)(b * c + a * d) * (long double)xd);
    R.m[1][1] = (float)((a * a + c * c - b * b - d * d) * yd);
    R.m[1][2] = (float)(2.0L * (long double)(c * d - a * b) * (long double)zd);
    R.m[2][0] = (float)(2.0L * (long double)(b * d - a * c) * (long double)xd);
    R.m[2][1] = (float)(2.0L * (long double)(c * d + a * b) * (long double)yd);
    R.m[2][2] = (float)((a * a + d * d - c * c - b * b) * zd);
    R.m[0][3] = qx;
    R.m[1][3] = qy;
    R.m[2][3] = qz;
    R.m[3][3] = 1.f;
    return R;
}

bool good_float(float v) { return std::isfinite(v); }

// niftilib's nifti_convert_nhdr2nim geometry plus reg_checkAndCorrectDimension.
void header_to_info(const Nifti1Header& h, nrh_nifti_info *o) {
    std::memset(o, 0, sizeof(*o));
    int dim[8];
    float pix[8];
    for (int i = 0; i < 8; ++i) {
        dim[i] = h.dim[i];
        pix[i] = h.pixdim[i];
    }
    for (int i = 2; i <= dim[0]; ++i)
        if (dim[i] <= 0) dim[i] = 1;
    for (int i = dim[0] + 1; i <= 7; ++i)
        if (dim[i] != 1 && dim[i] != 0) dim[i] = 1;
    for (int i = 1; i <= dim[0]; ++i)
        if (pix[i] == 0.0f || !good_float(pix[i])) pix[i] = 1.0f;
    // the qform: the quaternion when qform_code > 0; otherwise niftilib's diagonal of the
    // spacings, which reg_checkAndCorrectDimension rebuilds from the zero quaternion when no
    // sform is given either (the same matrix)
    const bool hasQ = h.qform_code > 0, hasS = h.sform_code > 0;
    mat44f qto{}, sto{};
    if (hasQ) {
        qto = quatern_to_mat44(h.quatern_b, h.quatern_c, h.quatern_d, h.qoffset_x, h.qoffset_y, h.qoffset_z, pix[1],
                               pix[2], pix[3], h.pixdim[0] < 0.0f ? -1.0f : 1.0f);
    } else {
        qto = quatern_to_mat44(0.f, 0.f, 0.f, 0.f, 0.f, 0.f, pix[1], pix[2], pix[3], 1.0f);
    }
    if (hasS) {
        for (int c = 0; c < 4; ++c) {
            sto.m[0][c] = h.srow_x[c];
            sto.m[1][c] = h.srow_y[c];
            sto.m[2][c] = h.srow_z[c];
        }
        sto.m[3][3] = 1.f;
    }
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            o->qto_xyz[r * 4 + c] = qto.m[r][c];
            o->sto_xyz[r * 4 + c] = sto.m[r][c];
        }
    // reg_checkAndCorrectDimension: no zero size, ndim = last axis > 1, slope 1 if unset,
    // spacings in mm (after the matrices, as there)
    for (int i = 1; i <= 7; ++i)
        if (dim[i] < 1) dim[i] = 1;
    for (int i = 1; i < 8; ++i)
        if (dim[i] > 1) dim[0] = i;
    const int units = h.xyzt_units & 0x07;
    if (units == 3) {  // micron
        for (int d = 1; d <= dim[0]; ++d) pix[d] /= 1000.f;
    } else if (units == 1) {  // metre
        for (int d = 1; d <= dim[0]; ++d) pix[d] *= 1000.f;
    }
    if (dim[2] == 1 && pix[2] == 0) pix[2] = 1;
    if (dim[3] == 1 && pix[3] == 0) pix[3] = 1;
    for (int i = 0; i < 8; ++i) {
        o->dim[i] = dim[i];
        o->pixdim[i] = pix[i];
    }
    o->datatype = h.datatype;
    o->scl_slope = h.scl_slope == 0.0f ? 1.0f : h.scl_slope;
    o->scl_inter = h.scl_inter;
    o->intent_code = h.intent_code;
    o->intent_p[0] = h.intent_p1;
    o->intent_p[1] = h.intent_p2;
    o->intent_p[2] = h.intent_p3;
    std::memcpy(o->intent_name, h.intent_name, 16);
    o->intent_name[15] = 0;
    std::memcpy(o->descrip, h.descrip, 80);
    o->descrip[79] = 0;
    o->cal_min = h.cal_min;
    o->cal_max = h.cal_max;
    o->xyzt_units = units == 0 ? 0 : 2;  // converted to mm
    o->xyzt_units |= h.xyzt_units & 0x38;
    o->qform_code = h.qform_code > 0 ? h.qform_code : 0;
    o->sform_code = h.sform_code > 0 ? h.sform_code : 0;
    o->quatern[0] = h.qform_code > 0 ? h.quatern_b : 0.f;
    o->quatern[1] = h.qform_code > 0 ? h.quatern_c : 0.f;
    o->quatern[2] = h.qform_code > 0 ? h.quatern_d : 0.f;
    o->qoffset[0] = h.qform_code > 0 ? h.qoffset_x : 0.f;
    o->qoffset[1] = h.qform_code > 0 ? h.qoffset_y : 0.f;
    o->qoffset[2] = h.qform_code > 0 ? h.qoffset_z : 0.f;
    o->qfac = (h.qform_code > 0 && h.pixdim[0] < 0.0f) ? -1.0f : 1.0f;
    for (int i = 0; i < 4; ++i) {
        o->srow[i] = h.srow_x[i];
        o->srow[4 + i] = h.srow_y[i];
        o->srow[8 + i] = h.srow_z[i];
    }
}

template<class T> void convert(const unsigned char *raw, size_t n, bool swapped, float *out) {
    for (size_t i = 0; i < n; ++i) {
        T v;
        std::memcpy(&v, raw + i * sizeof(T), sizeof(T));
        if (swapped) swap_bytes(v);
        out[i] = (float)v;
    }
}

bool ends_with(const std::string& s, const char *suffix) {
    const size_t n = std::strlen(suffix);
    return s.size() >= n && s.compare(s.size() - n, n, suffix) == 0;
}

void info_to_header(const nrh_nifti_info& o, Nifti1Header& h) {
    std::memset(&h, 0, sizeof(h));
    h.sizeof_hdr = 348;
    h.regular = 'r';
    for (int i = 0; i < 8; ++i) {
        h.dim[i] = (int16_t)o.dim[i];
        h.pixdim[i] = o.pixdim[i];
    }
    h.pixdim[0] = o.qfac < 0.0f ? -1.0f : 1.0f;
    h.intent_p1 = o.intent_p[0];
    h.intent_p2 = o.intent_p[1];
    h.intent_p3 = o.intent_p[2];
    h.intent_code = (int16_t)o.intent_code;
    h.datatype = 16;  // float32
    h.bitpix = 32;
    h.vox_offset = 352.f;
    h.scl_slope = o.scl_slope;
    h.scl_inter = o.scl_inter;
    h.xyzt_units = (char)o.xyzt_units;
    h.cal_max = o.cal_max;
    h.cal_min = o.cal_min;
    std::memcpy(h.descrip, o.descrip, 80);
    h.descrip[79] = 0;
    h.qform_code = (int16_t)o.qform_code;
    h.sform_code = (int16_t)o.sform_code;
    h.quatern_b = o.quatern[0];
    h.quatern_c = o.quatern[1];
    h.quatern_d = o.quatern[2];
    h.qoffset_x = o.qoffset[0];
    h.qoffset_y = o.qoffset[1];
    h.qoffset_z = o.qoffset[2];
    for (int i = 0; i < 4; ++i) {
        h.srow_x[i] = o.srow[i];
        h.srow_y[i] = o.srow[4 + i];
        h.srow_z[i] = o.srow[8 + i];
    }
    std::memcpy(h.intent_name, o.intent_name, 16);
    h.intent_name[15] = 0;
    std::memcpy(h.magic, "n+1", 4);
}

}  // namespace
}  // namespace nrh

using namespace nrh;

extern "C" {

int nrh_nifti_read_info(const char *path, nrh_nifti_info *info) {
    return guard([&] {
        NRH_REQUIRE(path && info, NRH_ERR_ARG, "null argument");
        GzFile gz;
        bool swapped;
        const Nifti1Header h = read_raw_header(path, gz, swapped);
        header_to_info(h, info);
    });
}

int nrh_nifti_read_data(const char *path, float *data, size_t capacity) {
    return guard([&] {
        NRH_REQUIRE(path && data, NRH_ERR_ARG, "null argument");
        GzFile gz;
        bool swapped;
        const Nifti1Header h = read_raw_header(path, gz, swapped);
        nrh_nifti_info info;
        header_to_info(h, &info);
        // the voxel count and byte size come from the file: multiply with overflow checks and
        // reject headers beyond 2^40 bytes of voxel data
        size_t n = 1;
        for (int i = 1; i <= 7; ++i) {
            const int d = (i > info.dim[0] && info.dim[i] == 0) ? 1 : info.dim[i];  // unused trailing axes
            NRH_REQUIRE(d >= 1, NRH_ERR_ARG, "bad image dimension in the header");
            NRH_REQUIRE(!__builtin_mul_overflow(n, (size_t)d, &n), NRH_ERR_ARG,
                        "image dimensions overflow the voxel count");
        }
        NRH_REQUIRE(capacity >= n, NRH_ERR_ARG, "output buffer smaller than the image");
        const int nb = bytes_of(h.datatype);
        size_t rawBytes = 0;
        NRH_REQUIRE(nb > 0 && !__builtin_mul_overflow(n, (size_t)nb, &rawBytes) && rawBytes <= ((size_t)1 << 40),
                    NRH_ERR_ARG, "image data size out of range");
        const long skip = (long)h.vox_offset - (long)sizeof(h);
        NRH_REQUIRE(skip >= 0, NRH_ERR_ARG, "bad vox_offset");
        if (skip > 0) NRH_REQUIRE(gzseek(gz.f, skip, SEEK_CUR) >= 0, NRH_ERR_ARG, "cannot seek to the voxel data");
        std::vector<unsigned char> raw(rawBytes);
        size_t got = 0;
        while (got < raw.size()) {
            const unsigned chunk = (unsigned)std::min<size_t>(raw.size() - got, 1u << 30);
            const int r = gzread(gz.f, raw.data() + got, chunk);
            NRH_REQUIRE(r > 0, NRH_ERR_ARG, std::string("truncated voxel data in ") + path);
            got += (size_t)r;
        }
        switch (h.datatype) {
        case 2: convert<uint8_t>(raw.data(), n, swapped, data); break;
        case 256: convert<int8_t>(raw.data(), n, swapped, data); break;
        case 4: convert<int16_t>(raw.data(), n, swapped, data); break;
        case 512: convert<uint16_t>(raw.data(), n, swapped, data); break;
        case 8: convert<int32_t>(raw.data(), n, swapped, data); break;
        case 768: convert<uint32_t>(raw.data(), n, swapped, data); break;
        case 16: convert<float>(raw.data(), n, swapped, data); break;
        case 64: convert<double>(raw.data(), n, swapped, data); break;
        case 1024: convert<int64_t>(raw.data(), n, swapped, data); break;
        case 1280: convert<uint64_t>(raw.data(), n, swapped, data); break;
        }
    });
}

int nrh_nifti_geometry(const nrh_nifti_info *info, nrh_geom *geom) {
    return guard([&] {
        NRH_REQUIRE(info && geom, NRH_ERR_ARG, "null argument");
        nrh_geom g{};
        g.nx = info->dim[1];
        g.ny = info->dim[2];
        g.nz = info->dim[3];
        g.dx = info->pixdim[1];
        g.dy = info->pixdim[2];
        g.dz = info->pixdim[3];
        std::memcpy(g.xyz, info->sform_code > 0 ? info->sto_xyz : info->qto_xyz, sizeof(g.xyz));
        mat44_inverse(g.xyz, g.ijk);
        g.sform = info->sform_code > 0 ? 1 : 0;
        *geom = g;
    });
}

int nrh_control_point_grid_info(const nrh_nifti_info *ref, const float spacing[3], nrh_nifti_info *grid) {
    return guard([&] {
        NRH_REQUIRE(ref && spacing && grid, NRH_ERR_ARG, "null argument");
        NRH_REQUIRE(ref->dim[3] > 1, NRH_ERR_UNSUPPORTED, "2-D images are not supported");
        // reg_createControlPointGrid (reg-lib/cpu/_reg_localTrans.cpp:21-128), header part
        auto ceil_int = [](float x) { const int i = (int)x; return i + (x > (float)i); };
        nrh_nifti_info g;
        std::memset(&g, 0, sizeof(g));
        g.dim[0] = 5;
        g.dim[1] = ceil_int((float)ref->dim[1] * ref->pixdim[1] / spacing[0] + 3.f);
        g.dim[2] = ceil_int((float)ref->dim[2] * ref->pixdim[2] / spacing[1] + 3.f);
        g.dim[3] = ceil_int((float)ref->dim[3] * ref->pixdim[3] / spacing[2] + 3.f);
        g.dim[4] = 1;
        g.dim[5] = 3;
        g.dim[6] = g.dim[7] = 1;
        g.pixdim[0] = 1.0f;
        g.pixdim[1] = spacing[0];
        g.pixdim[2] = spacing[1];
        g.pixdim[3] = spacing[2];
        for (int i = 4; i < 8; ++i) g.pixdim[i] = 1.0f;
        g.datatype = 16;
        g.scl_slope = 1.0f;
        g.xyzt_units = ref->xyzt_units;
        if (ref->qform_code + ref->sform_code > 0) {
            g.qform_code = ref->qform_code;
            g.sform_code = ref->sform_code;
        } else {
            g.qform_code = 1;
            g.sform_code = 0;
        }
        for (int i = 0; i < 3; ++i) g.quatern[i] = ref->quatern[i];
        g.qfac = ref->qfac;
        mat44f qto = quatern_to_mat44(ref->quatern[0], ref->quatern[1], ref->quatern[2], ref->qoffset[0],
                                      ref->qoffset[1], ref->qoffset[2], g.pixdim[1], g.pixdim[2], g.pixdim[3], g.qfac);
        // the origin moves by one node: -1 in every index
        auto mul = [](const mat44f& m, const float in[3], float out[3]) {
            for (int r = 0; r < 3; ++r)
                out[r] = (float)((double)m.m[r][0] * in[0] + (double)m.m[r][1] * in[1] + (double)m.m[r][2] * in[2] +
                                 (double)m.m[r][3]);
        };
        const float originIndex[3] = { -1.0f, -1.0f, -1.0f };
        float originReal[3];
        mul(qto, originIndex, originReal);
        for (int i = 0; i < 3; ++i) qto.m[i][3] = g.qoffset[i] = originReal[i];
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) g.qto_xyz[r * 4 + c] = qto.m[r][c];
        if (g.sform_code > 0) {
            const float ratio[3] = { g.pixdim[1] / ref->pixdim[1], g.pixdim[2] / ref->pixdim[2],
                                     g.pixdim[3] / ref->pixdim[3] };
            mat44f s{};
            for (int r = 0; r < 4; ++r)
                for (int c = 0; c < 4; ++c)
                    s.m[r][c] = (r < 3 && c < 3) ? ref->sto_xyz[r * 4 + c] * ratio[c] : ref->sto_xyz[r * 4 + c];
            mul(s, originIndex, originReal);
            for (int r = 0; r < 3; ++r) s.m[r][3] = originReal[r];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 4; ++c) g.srow[r * 4 + c] = s.m[r][c];
            for (int r = 0; r < 4; ++r)
                for (int c = 0; c < 4; ++c) g.sto_xyz[r * 4 + c] = s.m[r][c];
        }
        g.intent_code = 1007;  // NIFTI_INTENT_VECTOR
        std::strncpy(g.intent_name, "NREG_TRANS", 15);
        g.intent_p[0] = 2.0f;  // CUB_SPLINE_GRID (NREG_TRANS_TYPE, reg-lib/cpu/Maths.hpp:40-48)
        *grid = g;
    });
}

int nrh_nifti_write(const char *path, const nrh_nifti_info *info, const float *data) {
    return guard([&] {
        NRH_REQUIRE(path && info && data, NRH_ERR_ARG, "null argument");
        NRH_REQUIRE(info->dim[0] >= 1 && info->dim[0] <= 7, NRH_ERR_ARG, "bad dim[0]");
        size_t n = 1;
        for (int i = 1; i <= info->dim[0]; ++i) {
            NRH_REQUIRE(info->dim[i] >= 1 && info->dim[i] <= 32767, NRH_ERR_ARG, "bad image dimension");
            n *= (size_t)info->dim[i];
        }
        Nifti1Header h;
        info_to_header(*info, h);
        const std::string p(path);
        const bool gz = ends_with(p, ".gz");
        NRH_REQUIRE(ends_with(p, ".nii") || ends_with(p, ".nii.gz"), NRH_ERR_UNSUPPORTED,
                    "only single-file NIfTI-1 images (.nii, .nii.gz) are written");
        GzFile f;
        f.f = gzopen(path, gz ? "wb6" : "wbT");
        NRH_REQUIRE(f.f, NRH_ERR_ARG, std::string("cannot create ") + path);
        const char ext[4] = { 0, 0, 0, 0 };
        NRH_REQUIRE(gzwrite(f.f, &h, sizeof(h)) == (int)sizeof(h) && gzwrite(f.f, ext, 4) == 4, NRH_ERR_ARG,
                    "write failed");
        const unsigned char *p8 = reinterpret_cast<const unsigned char *>(data);
        size_t left = n * sizeof(float);
        while (left > 0) {
            const unsigned chunk = (unsigned)std::min<size_t>(left, 1u << 30);
            NRH_REQUIRE(gzwrite(f.f, p8, chunk) == (int)chunk, NRH_ERR_ARG, "write failed");
            p8 += chunk;
            left -= chunk;
        }
        const int rc = gzclose(f.f);
        f.f = nullptr;
        NRH_REQUIRE(rc == Z_OK, NRH_ERR_ARG, "close failed");
    });
}

int nrh_affine_read(const char *path, float m[16]) {
    return guard([&] {
        NRH_REQUIRE(path && m, NRH_ERR_ARG, "null argument");
        FILE *f = std::fopen(path, "r");
        NRH_REQUIRE(f, NRH_ERR_ARG, std::string("The affine file can not be read: ") + path);
        double v[16];
        int got = 0;
        for (; got < 16; ++got)
            if (std::fscanf(f, "%lf", &v[got]) != 1) break;
        std::fclose(f);
        NRH_REQUIRE(got == 16, NRH_ERR_ARG, std::string("an affine file holds 16 numbers: ") + path);
        for (int i = 0; i < 16; ++i) m[i] = (float)v[i];
    });
}

int nrh_affine_write(const char *path, const float m[16]) {
    return guard([&] {
        NRH_REQUIRE(path && m, NRH_ERR_ARG, "null argument");
        FILE *f = std::fopen(path, "w");
        NRH_REQUIRE(f, NRH_ERR_ARG, std::string("cannot create ") + path);
        for (int r = 0; r < 4; ++r)
            std::fprintf(f, "%.7g %.7g %.7g %.7g\n", (double)m[r * 4], (double)m[r * 4 + 1], (double)m[r * 4 + 2],
                         (double)m[r * 4 + 3]);
        NRH_REQUIRE(std::fclose(f) == 0, NRH_ERR_ARG, "close failed");
    });
}

}  // extern "C"
