// iq_prep.hip -- host preparation of the tile loop's inputs, inside the C ABI
// (utils.jl:69-129): element-type conversion with missing values, soft-data
// padding, and the disabled-candidate map.  Host code only; included by the
// unity build in iq_api.hip.

namespace {

size_t raw_size(int t) {
  switch (t) {
    case IQ_RAW_F32: case IQ_RAW_I32: case IQ_RAW_U32: return 4;
    case IQ_RAW_F64: case IQ_RAW_I64: case IQ_RAW_U64: return 8;
    case IQ_RAW_U8: case IQ_RAW_I8: return 1;
    case IQ_RAW_U16: case IQ_RAW_I16: return 2;
    default: return 0;
  }
}

// Value i of a raw array as a double, and whether it is inactive (missing or NaN).
inline double raw_get(const void* p, int t, int64_t i, const uint8_t* missing, bool* inactive) {
  double v;
  switch (t) {
    case IQ_RAW_F32: v = (double)((const float*)p)[i]; break;
    case IQ_RAW_F64: v = ((const double*)p)[i]; break;
    case IQ_RAW_U8: v = (double)((const uint8_t*)p)[i]; break;
    case IQ_RAW_I8: v = (double)((const int8_t*)p)[i]; break;
    case IQ_RAW_U16: v = (double)((const uint16_t*)p)[i]; break;
    case IQ_RAW_I16: v = (double)((const int16_t*)p)[i]; break;
    case IQ_RAW_U32: v = (double)((const uint32_t*)p)[i]; break;
    case IQ_RAW_I32: v = (double)((const int32_t*)p)[i]; break;
    case IQ_RAW_U64: v = (double)((const uint64_t*)p)[i]; break;
    default: v = (double)((const int64_t*)p)[i]; break;
  }
  *inactive = (missing && missing[i]) || std::isnan(v);
  return v;
}

void check_raw(const void* p, int t, const int64_t* dims, int64_t* n) {
  if (!p || !dims) bad(IQ_ERR_ARG, "null argument");
  if (raw_size(t) == 0) bad(IQ_ERR_ARG, "unknown raw element type");
  *n = 1;
  for (int d = 0; d < 3; d++) {
    if (dims[d] <= 0) bad(IQ_ERR_ARG, "invalid extents");
    *n *= dims[d];
  }
}

// ImageFiltering Pad(:symmetric) source index of position i >= 0 along an
// axis of n voxels: mirror with the edge voxel repeated, period 2n.
inline int64_t sym_index(int64_t i, int64_t n) {
  int64_t k = i % (2 * n);
  return k < n ? k : 2 * n - 1 - k;
}

}  // namespace

extern "C" {

int iq_prepare(const void* raw, int raw_dtype, const uint8_t* missing, const int64_t dims[3], void* out) {
  return guard(nullptr, [&] {
    int64_t n;
    check_raw(raw, raw_dtype, dims, &n);
    if (!out) bad(IQ_ERR_ARG, "null output");
    if (raw_dtype == IQ_RAW_F32) {
      float* o = (float*)out;
      for (int64_t i = 0; i < n; i++) {
        bool off;
        float v = ((const float*)raw)[i];
        raw_get(raw, raw_dtype, i, missing, &off);
        o[i] = off ? 0.0f : v;
      }
    } else {
      double* o = (double*)out;
      for (int64_t i = 0; i < n; i++) {
        bool off;
        double v = raw_get(raw, raw_dtype, i, missing, &off);
        o[i] = off ? 0.0 : v;
      }
    }
  });
}

int iq_prepare_soft(const void* aux, int raw_dtype, const uint8_t* missing, const int64_t dims[3],
                    const int64_t padsize[3], double* out) {
  return guard(nullptr, [&] {
    int64_t n;
    check_raw(aux, raw_dtype, dims, &n);
    if (!out || !padsize) bad(IQ_ERR_ARG, "null argument");
    for (int d = 0; d < 3; d++)
      if (padsize[d] <= 0) bad(IQ_ERR_ARG, "invalid padsize");
    for (int64_t z = 0; z < padsize[2]; z++) {
      const int64_t sz = sym_index(z, dims[2]);
      for (int64_t y = 0; y < padsize[1]; y++) {
        const int64_t sy = sym_index(y, dims[1]);
        for (int64_t x = 0; x < padsize[0]; x++) {
          const int64_t sx = sym_index(x, dims[0]);
          bool off;
          double v = raw_get(aux, raw_dtype, sx + dims[0] * (sy + dims[1] * sz), missing, &off);
          out[x + padsize[0] * (y + padsize[1] * z)] = off ? 0.0 : v;
        }
      }
    }
  });
}

int iq_finddisabled(const void* raw_ti, int raw_dtype, const uint8_t* missing, const int64_t tisize[3],
                    const int64_t tilesize[3], uint8_t* disabled) {
  return guard(nullptr, [&] {
    int64_t n;
    check_raw(raw_ti, raw_dtype, tisize, &n);
    if (!tilesize || !disabled) bad(IQ_ERR_ARG, "null argument");
    int64_t ds[3];
    for (int d = 0; d < 3; d++) {
      ds[d] = tisize[d] - tilesize[d] + 1;
      if (tilesize[d] <= 0 || ds[d] <= 0) bad(IQ_ERR_ARG, "tile larger than the training image");
    }
    // inclusive 3-D prefix counts of inactive voxels, one guard layer per axis
    const int64_t X = tisize[0] + 1, Y = tisize[1] + 1, Z = tisize[2] + 1;
    std::vector<int64_t> S((size_t)(X * Y * Z), 0);
    auto at = [&](int64_t x, int64_t y, int64_t z) -> int64_t& { return S[(size_t)(x + X * (y + Y * z))]; };
    for (int64_t z = 1; z < Z; z++)
      for (int64_t y = 1; y < Y; y++)
        for (int64_t x = 1; x < X; x++) {
          bool off;
          raw_get(raw_ti, raw_dtype, (x - 1) + tisize[0] * ((y - 1) + tisize[1] * (z - 1)), missing, &off);
          at(x, y, z) = (off ? 1 : 0) + at(x - 1, y, z) + at(x, y - 1, z) + at(x, y, z - 1) - at(x - 1, y - 1, z) -
                        at(x - 1, y, z - 1) - at(x, y - 1, z - 1) + at(x - 1, y - 1, z - 1);
        }
    const int64_t tx = tilesize[0], ty = tilesize[1], tz = tilesize[2];
    for (int64_t z = 0; z < ds[2]; z++)
      for (int64_t y = 0; y < ds[1]; y++)
        for (int64_t x = 0; x < ds[0]; x++) {
          int64_t c = at(x + tx, y + ty, z + tz) - at(x, y + ty, z + tz) - at(x + tx, y, z + tz) -
                      at(x + tx, y + ty, z) + at(x, y, z + tz) + at(x, y + ty, z) + at(x + tx, y, z) - at(x, y, z);
          disabled[x + ds[0] * (y + ds[1] * z)] = c > 0;
        }
  });
}

}  // extern "C"
