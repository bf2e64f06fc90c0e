/* fdt_io.h -- NIfTI-1 images and mask maps of the in-process bedpostx_gpu
 * pipeline (fdt_amd/bin/bedpostx_gpu), as a C ABI.
 *
 * They replace the FSL image calls the reference pipeline makes through its
 * tools: read_volume4D + volume4D::matrix(mask) in split_parts_gpu
 * (CUDA/split_parts_gpu.cc:96-119) and volume4D::setmatrix + save_volume4D in
 * merge_parts_gpu (merge_parts_gpu.cc:74-81).  Images are single-file NIfTI-1
 * (.nii or .nii.gz, either byte order, integer or float voxels) read as
 * float32; written images are float32 with the spatial header of a template.
 * Mask voxels are ordered as FSL orders them: x fastest, then y, then z.
 */
#ifndef FDT_IO_H
#define FDT_IO_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fdt_image fdt_image;

/* Read `path` (FSL name resolution: path, path.nii.gz, path.nii).  NULL on
 * failure, with the reason in err[errlen]. */
fdt_image* fdt_image_read(const char* path, char* err, int errlen);
void fdt_image_free(fdt_image* im);
/* dim[8] (dim[0] = rank) and pixdim[8] of the header */
void fdt_image_dims(const fdt_image* im, int* dim, float* pixdim);
/* the voxels, x fastest, volume after volume */
const float* fdt_image_data(const fdt_image* im);
/* Write nt volumes `data` (x fastest) with the grid, voxel size, units and
 * qform/sform of `like`; cal_min/cal_max are the display range. */
int fdt_image_write(const char* path, const fdt_image* like, int nt, const float* data, float cal_min, float cal_max,
                    char* err, int errlen);
/* number of voxels where the first volume of `mask` is non-zero */
long fdt_image_mask_count(const fdt_image* mask);
/* out[nvox][nt]: `img` at the mask voxels (volume4D::matrix); returns nvox,
 * or -1 when the grids differ */
long fdt_image_mask_matrix(const fdt_image* img, const fdt_image* mask, float* out);

#ifdef __cplusplus
}
#endif
#endif
