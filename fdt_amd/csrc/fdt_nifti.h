// fdt_nifti.h -- NIfTI-1 single-file images (.nii, .nii.gz) as float32 volumes,
// and the mask <-> voxel-matrix maps of the bedpostx pipeline.
//
// The pipeline reads data, nodif_brain_mask and grad_dev and writes the
// merged sample volumes, means and dyads that FSL's volume4D::matrix(mask) /
// setmatrix(matrix, mask) produce (split_parts_gpu.cc:104, merge_parts_gpu.cc:77):
// mask voxels in storage order (x fastest, then y, then z).  Format per the
// public NIfTI-1 specification: a 348-byte header, a 4-byte extension flag,
// data at vox_offset; either byte order is read, little-endian is written.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace fdt {

struct Nifti {
  int dim[8] = {0};        // dim[0] = rank, dim[1..7] extents
  float pixdim[8] = {0};
  int datatype = 16;       // of the source file
  float scl_slope = 0.0f, scl_inter = 0.0f;
  float cal_max = 0.0f, cal_min = 0.0f;
  uint8_t xyzt_units = 0;
  int16_t qform_code = 0, sform_code = 0;
  float quatern[6] = {0};  // b, c, d, qoffset x, y, z
  float srow[12] = {0};    // srow_x, srow_y, srow_z
  std::vector<float> data; // nx*ny*nz*nt values, x fastest

  size_t nxyz() const { return (size_t)dim[1] * (dim[2] > 0 ? dim[2] : 1) * (dim[3] > 0 ? dim[3] : 1); }
  int nt() const { return dim[0] >= 4 && dim[4] > 0 ? dim[4] : 1; }
};

// FSL's image name resolution (imtest / read_volume): `base` itself, then
// base.nii.gz, base.nii.  Empty if none exists.
std::string image_path(const std::string& base);

// Read an image as float32 (scl_slope / scl_inter applied when slope != 0).
bool read_nifti(const std::string& path, Nifti& img, std::string& err);

// Write `nt` volumes of float32 with the spatial header (dims 1..3, pixdim,
// units, qform, sform) of `like`; .gz paths are gzip-compressed (level 1, on
// `threads` threads).
bool write_nifti(const std::string& path, const Nifti& like, int nt, const float* data, float cal_min, float cal_max,
                 std::string& err, int threads = 1);

// Indices (into one volume) of the voxels where mask != 0, in storage order.
std::vector<size_t> mask_indices(const Nifti& mask);

// [nvox][nt] voxel-major matrix of `img` at the mask voxels (volume4D::matrix).
void mask_matrix(const Nifti& img, const std::vector<size_t>& idx, float* out);

// Scatter a [nt][nvox] (volume-major) matrix into nt volumes of nxyz values,
// zero outside the mask (volume4D::setmatrix).
void scatter_volumes(const std::vector<size_t>& idx, size_t nxyz, int nt, const float* m, float* vol);

}  // namespace fdt
