// Output files of the renders: the FITS spectral cube (NOST 100-2.0 primary
// HDU, the header cards and data order of src/fits_writer.cpp:148-401) and
// the Radiance RGBE .hdr image (the CLI's linear and exposure-scaled HDR
// outputs, cli/main.cpp:398-434).
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace grrt_amd {

// Metadata cards of a spectral cube (src/fits_writer.cpp:210-224).
struct FitsMeta {
    double spin = 0.0, mass = 1.0, observer_r = 50.0, observer_theta = 1.396, fov = 1.047;
    int samples_per_pixel = 1;
};

// The header as 80-character cards, END included (not yet padded to 2880).
// bitpix -64 (double) or -32 (float).
std::vector<std::string> fits_header_cards(int width, int height, int num_bins, const std::vector<double>& freqs,
                                           const FitsMeta& meta, int bitpix);

// Row-streamed cube writer: the header is written and the file extended to
// its final padded size at open; write_row() stores one image row (renderer
// row j, 0 = top, layout row[i * num_bins + k]) at its place in every
// frequency plane, bottom row first as FITS counts rows.  Rows may arrive in
// any order; unwritten rows read as zero.
class FitsCubeWriter {
public:
    FitsCubeWriter(const std::string& path, int width, int height, int num_bins, const std::vector<double>& freqs,
                   const FitsMeta& meta, int bitpix);
    ~FitsCubeWriter();
    FitsCubeWriter(const FitsCubeWriter&) = delete;
    FitsCubeWriter& operator=(const FitsCubeWriter&) = delete;
    void write_row(int j, const double* row);
    void write_row(int j, const float* row);
    void finish();

private:
    template <class T>
    void put_row(int j, const T* row);
    int fd_ = -1;
    int width_, height_, bins_, bitpix_;
    long long data_start_ = 0;
    std::string path_;
    std::vector<unsigned char> line_;
};

// Row-streamed Radiance RGBE image (new-style run-length scanlines), top
// row first ("-Y height +X width"): the header is written at open and rows
// are appended in order as they arrive.  finish() writes any rows not yet
// written as black, so the file always holds `height` rows.
class RgbeWriter {
public:
    RgbeWriter(const std::string& path, int width, int height);
    ~RgbeWriter();
    RgbeWriter(const RgbeWriter&) = delete;
    RgbeWriter& operator=(const RgbeWriter&) = delete;
    // The next `rows` rows (rgba: rows*width*4 floats); each RGB is scaled by
    // `scale` before encoding.
    void write_rows(const float* rgba, int rows, float scale);
    void finish();
    int rows_written() const { return next_; }

private:
    std::FILE* f_ = nullptr;
    int width_, height_, next_ = 0;
    std::string path_;
    std::vector<unsigned char> px_, out_;
};

// The whole frame through RgbeWriter.  rgba: width*height*4 floats.
void write_rgbe(const std::string& path, const float* rgba, int width, int height, float scale);

// Exposure statistics of the CLI's normalised HDR (cli/main.cpp:410-425),
// accumulated over pixels in index order: 0.18 over the log-average
// luminance of the lit pixels.  Feeding a frame's rows in order gives the
// same sum, term by term, as one pass over the frame.
struct ExposureSum {
    double log_sum = 0.0;
    long long lit = 0;
    void add(const float* rgba, long long n_pixels);
    float scale() const;
};

float hdr_exposure_scale(const float* rgba, int width, int height);

// An equirectangular background image for GRRT_BG_TEXTURE: linear RGB
// floats, row 0 at the top (theta = 0), column 0 at phi = -pi.
struct Texture {
    int width = 0, height = 0;
    std::vector<float> rgb;   // width * height * 3
};

// Reads a Radiance RGBE .hdr ("-Y height +X width"; flat or new-style
// run-length scanlines; value = (mantissa + 0.5) * 2^(exponent - 136), zero for exponent 0) or a PFM
// (colour "PF" or grey "Pf", byte order from the scale's sign, rows stored
// bottom first).  Throws std::runtime_error with the reason on anything
// else.
Texture read_texture(const std::string& path);

}  // namespace grrt_amd
