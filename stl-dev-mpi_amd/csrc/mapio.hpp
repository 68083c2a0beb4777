// Map I/O of the statistics and synthesis tools: full-sky NESTED maps as
// `.npy` (any float / integer dtype, either byte order) or HEALPix FITS
// binary tables (RING maps are reordered to NESTED on reading), read whole or
// by pixel ranges and written whole or by slices; and the counter-based
// standard normal draws of the initial maps.  Host code, no device work.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace stl {
namespace io {

// nside of a full-sky map of npix pixels (throws otherwise).
int64_t nside_of(size_t npix);

// NESTED index of RING pixel pix at the given order (log2 nside).
int64_t ring2nest(int64_t pix, int order);

// A `.npy` array as double with its shape (C order).
std::vector<double> read_npy(const std::string& path, std::vector<size_t>* shape);
// Little-endian float64 `.npy` (format 1.0) of the given shape.
void write_npy(const std::string& path, const double* values, const std::vector<size_t>& shape);

// Full-sky NESTED map as a HEALPix binary table, one float32 (`E`) or
// float64 (`D`) column.
void write_healpix_fits(const std::string& path, const double* values, size_t n, const std::string& column,
                        bool dbl);
// First column of the first binary table, as a NESTED map.
std::vector<double> read_healpix_fits(const std::string& path);

// By extension: `.fits` / `.fit` as FITS, anything else as `.npy`.
std::vector<double> read_map(const std::string& path);
size_t map_len(const std::string& path);
std::vector<double> read_map_range(const std::string& path, size_t start, size_t end);

// A map file written by slices: create once (headers, file sized for n
// pixels), then write pixel ranges at their offsets.
void sliced_create(const std::string& path, size_t n, const std::string& column, bool dbl);
void sliced_write(const std::string& path, bool dbl, size_t start, const double* values, size_t n);

// Standard normal draw of pixel id for seed: splitmix64 of the seed and of
// the id, then Box-Muller on two 53-bit uniforms; the same map whatever the
// decomposition.
double pixel_normal(uint64_t seed, int64_t id);

}  // namespace io
}  // namespace stl
