// nifti_io.h -- single-file NIfTI-1 volumes (.nii, .nii.gz) for the
// RandomiseGroupLevel front end.  Reads the voxel types the reference CLI
// accepts (code/Bash_Wrapper/RandomiseGroupLevel.cpp:1306-1395: uint8,
// int16, uint16, float32; int32 and float64 are also read) into float, in
// either byte order; writes float32 volumes with the header of an input
// image (code/Bash_Wrapper/HelpFunctions.cpp:325-406).  The scaling fields
// (scl_slope / scl_inter) are carried over, not applied, as in the
// reference, which converts the raw stored values.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace bpg_cli {

// The 348-byte NIfTI-1 header, field for field.
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
    float slice_duration;
    float toffset;
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

enum : int16_t {
    NIFTI_UINT8 = 2, NIFTI_INT16 = 4, NIFTI_INT32 = 8, NIFTI_FLOAT32 = 16, NIFTI_FLOAT64 = 64,
    NIFTI_UINT16 = 512
};

struct Volume {
    Nifti1Header hdr{};     // as read (native byte order after reading)
    std::string fname;      // the name it was read from
    size_t nx = 1, ny = 1, nz = 1, nt = 1;
    std::vector<float> data;  // nx*ny*nz*nt, x fastest
};

// Returns false with a message in err when the file cannot be read or has a
// voxel type other than those listed above.
bool read_volume(const std::string& path, Volume& out, std::string& err);

// Writes nt volumes of float data with the header of `like` (dims 1-4 set
// from like.nx..nz and nt, datatype float32, cal_min/cal_max set to the data
// range as HelpFunctions.cpp:379-381).  .nii.gz names are compressed.
bool write_volume(const std::string& path, const Volume& like, size_t nt, const float* data, std::string& err);

// The first '.' of the file name part (after the last '/').  The reference
// takes the first '.' of the whole path (HelpFunctions.cpp:6-29, 341-367),
// which is the same for names without a directory and would otherwise put
// the suffix into a directory name such as "./data".
size_t first_dot(const std::string& name);

// true for names ending in .nii or .nii.gz after that dot, the test of
// CheckFileExtension (HelpFunctions.cpp:6-29)
bool extension_ok(const std::string& name, std::string& ext);

// name with suffix inserted before that dot (HelpFunctions.cpp:341-367)
std::string insert_suffix(const std::string& name, const std::string& suffix);

}  // namespace bpg_cli
