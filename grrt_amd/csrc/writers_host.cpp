// FITS cube and Radiance RGBE writers (host C++).
#include "writers_host.hpp"

#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>

namespace grrt_amd {
namespace {

constexpr int kCard = 80;
constexpr int kBlock = 2880;

// One header card: keyword in columns 1-8, "= ", the value right-justified
// in columns 11-30, then " / comment"; blank-padded (or cut) to 80 columns.
std::string card(const std::string& keyword, std::string value, const std::string& comment) {
    std::string s = keyword;
    s.resize(8, ' ');
    s += "= ";
    if (value.size() < 20) value.insert(0, 20 - value.size(), ' ');
    s += value;
    if (!comment.empty()) s += " / " + comment;
    s.resize(kCard, ' ');
    return s;
}

std::string fmt_double(double v) {
    char b[64];
    std::snprintf(b, sizeof b, "%.12E", v);
    return b;
}

std::string fmt_int(long long v) { return std::to_string(v); }

// Quoted string value: at least 8 characters inside the quotes, at most 18.
std::string fmt_string(std::string v) {
    if (v.size() < 8) v.resize(8, ' ');
    if (v.size() > 18) v.resize(18);
    return "'" + v + "'";
}

bool uniform_linear(const std::vector<double>& f) {
    if (f.size() < 2) return true;
    const double step = f[1] - f[0];
    if (step == 0.0) return false;
    for (size_t i = 2; i < f.size(); ++i)
        if (std::fabs(((f[i] - f[i - 1]) - step) / step) > 1e-4) return false;
    return true;
}

bool uniform_log(const std::vector<double>& f) {
    if (f.size() < 2) return true;
    if (f[0] <= 0.0 || f[1] <= 0.0) return false;
    const double step = std::log10(f[1]) - std::log10(f[0]);
    if (step == 0.0) return false;
    for (size_t i = 2; i < f.size(); ++i) {
        if (f[i] <= 0.0) return false;
        if (std::fabs(((std::log10(f[i]) - std::log10(f[i - 1])) - step) / step) > 1e-4) return false;
    }
    return true;
}

std::string freq_key(int k) {
    char b[16];
    std::snprintf(b, sizeof b, "FREQ%04d", k);
    return b;
}

void put_be(unsigned char* dst, double v) {
    uint64_t u;
    std::memcpy(&u, &v, 8);
    for (int b = 0; b < 8; ++b) dst[b] = static_cast<unsigned char>(u >> (56 - 8 * b));
}

void put_be(unsigned char* dst, float v) {
    uint32_t u;
    std::memcpy(&u, &v, 4);
    for (int b = 0; b < 4; ++b) dst[b] = static_cast<unsigned char>(u >> (24 - 8 * b));
}

void write_all(int fd, const void* p, size_t n, long long off, const std::string& path) {
    const char* c = static_cast<const char*>(p);
    while (n > 0) {
        const ssize_t w = ::pwrite(fd, c, n, static_cast<off_t>(off));
        if (w < 0) {
            if (errno == EINTR) continue;
            throw std::runtime_error("write error on '" + path + "': " + std::strerror(errno));
        }
        c += w;
        off += w;
        n -= static_cast<size_t>(w);
    }
}

}  // namespace

std::vector<std::string> fits_header_cards(int width, int height, int num_bins, const std::vector<double>& freqs,
                                           const FitsMeta& meta, int bitpix) {
    std::vector<std::string> c;
    c.push_back(card("SIMPLE", "T", "conforms to FITS standard"));
    c.push_back(card("BITPIX", fmt_int(bitpix), bitpix == -64 ? "IEEE 754 double precision"
                                                              : "IEEE 754 single precision"));
    c.push_back(card("NAXIS", "3", "number of array dimensions"));
    c.push_back(card("NAXIS1", fmt_int(width), "pixels along RA axis"));
    c.push_back(card("NAXIS2", fmt_int(height), "pixels along Dec axis"));
    c.push_back(card("NAXIS3", fmt_int(num_bins), "number of frequency bins"));
    c.push_back(card("BUNIT", fmt_string("erg/s/cm2/Hz/sr"), "specific intensity"));
    if (!freqs.empty()) {
        const int nk = std::min<int>(num_bins, static_cast<int>(freqs.size()));
        if (freqs.size() >= 2 && uniform_linear(freqs)) {
            c.push_back(card("CTYPE3", fmt_string("FREQ"), "frequency"));
            c.push_back(card("CUNIT3", fmt_string("Hz"), "frequency unit"));
            c.push_back(card("CRPIX3", fmt_double(1.0), "reference pixel"));
            c.push_back(card("CRVAL3", fmt_double(freqs[0]), "frequency at ref pixel (Hz)"));
            c.push_back(card("CDELT3", fmt_double(freqs[1] - freqs[0]), "frequency step (Hz)"));
        } else {
            const bool lg = freqs.size() >= 2 && uniform_log(freqs);
            if (lg) {
                c.push_back(card("CTYPE3", fmt_string("FREQ-LOG"), "log10(frequency/Hz)"));
                c.push_back(card("CUNIT3", fmt_string("log(Hz)"), "log10 of frequency in Hz"));
                c.push_back(card("CRPIX3", fmt_double(1.0), "reference pixel"));
                c.push_back(card("CRVAL3", fmt_double(std::log10(freqs[0])), "log10(freq) at ref pixel"));
                c.push_back(card("CDELT3", fmt_double(std::log10(freqs[1]) - std::log10(freqs[0])),
                                 "log10(freq) step per pixel"));
            } else {
                c.push_back(card("CTYPE3", fmt_string("FREQ-TAB"), "frequency (see FREQnnnn keys)"));
                c.push_back(card("CUNIT3", fmt_string("Hz"), "frequency unit"));
                c.push_back(card("CRPIX3", fmt_double(1.0), "reference pixel"));
                c.push_back(card("CRVAL3", fmt_double(1.0), "slice index"));
                c.push_back(card("CDELT3", fmt_double(1.0), "slice index step"));
            }
            for (int k = 0; k < nk; ++k)
                c.push_back(card(freq_key(k + 1), fmt_double(freqs[k]), "freq bin " + std::to_string(k + 1) + " (Hz)"));
        }
    }
    c.push_back(card("SPIN", fmt_double(meta.spin), "Kerr spin parameter a/M"));
    c.push_back(card("MASS", fmt_double(meta.mass), "black hole mass (geometrized units)"));
    c.push_back(card("OBS_R", fmt_double(meta.observer_r), "observer radial coordinate (M)"));
    c.push_back(card("OBS_TH", fmt_double(meta.observer_theta), "observer polar angle (rad)"));
    c.push_back(card("FOV", fmt_double(meta.fov), "camera field of view (rad)"));
    c.push_back(card("SPP", fmt_int(meta.samples_per_pixel), "samples per pixel (anti-aliasing)"));
    c.push_back(card("ORIGIN", fmt_string("grrt"), "GR ray tracer"));
    std::string end = "END";
    end.resize(kCard, ' ');
    c.push_back(end);
    return c;
}

FitsCubeWriter::FitsCubeWriter(const std::string& path, int width, int height, int num_bins,
                               const std::vector<double>& freqs, const FitsMeta& meta, int bitpix)
    : width_(width), height_(height), bins_(num_bins), bitpix_(bitpix), path_(path) {
    if (width <= 0 || height <= 0 || num_bins <= 0) throw std::runtime_error("FITS cube with an empty axis");
    if (bitpix != -64 && bitpix != -32) throw std::runtime_error("FITS BITPIX must be -64 or -32");
    fd_ = ::open(path.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
    if (fd_ < 0) throw std::runtime_error("cannot open '" + path + "' for writing: " + std::strerror(errno));
    const auto cards = fits_header_cards(width, height, num_bins, freqs, meta, bitpix);
    std::string header;
    for (const auto& c : cards) header += c;
    header.resize((header.size() + kBlock - 1) / kBlock * kBlock, ' ');
    write_all(fd_, header.data(), header.size(), 0, path_);
    data_start_ = static_cast<long long>(header.size());
    const long long bytes = static_cast<long long>(width) * height * num_bins * (bitpix == -64 ? 8 : 4);
    const long long padded = (bytes + kBlock - 1) / kBlock * kBlock;
    if (::ftruncate(fd_, static_cast<off_t>(data_start_ + padded)) != 0)
        throw std::runtime_error("cannot extend '" + path + "': " + std::strerror(errno));
    line_.resize(static_cast<size_t>(width) * (bitpix == -64 ? 8 : 4));
}

FitsCubeWriter::~FitsCubeWriter() {
    if (fd_ >= 0) ::close(fd_);
}

template <class T>
void FitsCubeWriter::put_row(int j, const T* row) {
    if (j < 0 || j >= height_) throw std::runtime_error("FITS row out of range");
    const int fits_row = height_ - 1 - j;        // FITS pixel (1,1) is bottom-left
    const size_t es = bitpix_ == -64 ? 8 : 4;
    for (int k = 0; k < bins_; ++k) {
        for (int i = 0; i < width_; ++i) {
            const T v = row[static_cast<size_t>(i) * bins_ + k];
            if (es == 8) put_be(&line_[i * es], static_cast<double>(v));
            else put_be(&line_[i * es], static_cast<float>(v));
        }
        const long long off =
            data_start_ + (static_cast<long long>(k) * height_ * width_ + static_cast<long long>(fits_row) * width_) *
                              static_cast<long long>(es);
        write_all(fd_, line_.data(), line_.size(), off, path_);
    }
}

void FitsCubeWriter::write_row(int j, const double* row) { put_row(j, row); }
void FitsCubeWriter::write_row(int j, const float* row) { put_row(j, row); }

void FitsCubeWriter::finish() {
    if (fd_ < 0) return;
    if (::fsync(fd_) != 0 && errno != EINVAL) throw std::runtime_error("flush error on '" + path_ + "'");
    if (::close(fd_) != 0) throw std::runtime_error("close error on '" + path_ + "'");
    fd_ = -1;
}

RgbeWriter::RgbeWriter(const std::string& path, int width, int height) : width_(width), height_(height), path_(path) {
    if (width <= 0 || height <= 0) throw std::runtime_error("empty image");
    f_ = std::fopen(path.c_str(), "wb");
    if (!f_) throw std::runtime_error("cannot open '" + path + "' for writing: " + std::strerror(errno));
    if (std::fprintf(f_, "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n", height, width) < 0) {
        std::fclose(f_);
        f_ = nullptr;
        throw std::runtime_error("write error on '" + path + "'");
    }
    px_.resize(static_cast<size_t>(width) * 4);
    out_.reserve(static_cast<size_t>(width) * 5 + 4);
}

RgbeWriter::~RgbeWriter() {
    if (f_) std::fclose(f_);
}

void RgbeWriter::write_rows(const float* rgba, int rows, float scale) {
    if (!f_) throw std::runtime_error("RGBE writer is closed");
    if (rows < 0 || next_ + rows > height_) throw std::runtime_error("RGBE rows past the image height");
    const int width = width_;
    for (int j = 0; j < rows; ++j, ++next_) {
        for (int i = 0; i < width; ++i) {
            unsigned char* q = &px_[static_cast<size_t>(i) * 4];
            if (!rgba) {
                q[0] = q[1] = q[2] = q[3] = 0;
                continue;
            }
            const float* p = rgba + (static_cast<size_t>(j) * width + i) * 4;
            float c[3];
            for (int k = 0; k < 3; ++k) {
                const float v = p[k] * scale;
                c[k] = (v > 0.0f && std::isfinite(v)) ? v : 0.0f;   // RGBE holds non-negative finite values
            }
            const float m = std::max(c[0], std::max(c[1], c[2]));
            if (m < 1e-32f) {
                q[0] = q[1] = q[2] = q[3] = 0;
            } else {
                int e;
                const float s = std::frexp(m, &e) * 256.0f / m;   // Ward's shared-exponent encoding
                for (int k = 0; k < 3; ++k) q[k] = static_cast<unsigned char>(std::min(255.0f, c[k] * s));
                q[3] = static_cast<unsigned char>(e + 128);
            }
        }
        out_.clear();
        if (width >= 8 && width < 32768) {
            // New-style scanline: marker, then each channel as literal runs of <= 128 bytes.
            out_.push_back(2);
            out_.push_back(2);
            out_.push_back(static_cast<unsigned char>(width >> 8));
            out_.push_back(static_cast<unsigned char>(width & 255));
            for (int ch = 0; ch < 4; ++ch)
                for (int i = 0; i < width;) {
                    const int n = std::min(128, width - i);
                    out_.push_back(static_cast<unsigned char>(n));
                    for (int t = 0; t < n; ++t) out_.push_back(px_[static_cast<size_t>(i + t) * 4 + ch]);
                    i += n;
                }
        } else {
            out_.assign(px_.begin(), px_.end());
        }
        if (std::fwrite(out_.data(), 1, out_.size(), f_) != out_.size())
            throw std::runtime_error("write error on '" + path_ + "'");
    }
}

void RgbeWriter::finish() {
    if (!f_) return;
    if (next_ < height_) write_rows(nullptr, height_ - next_, 1.0f);
    const int rc = std::fclose(f_);
    f_ = nullptr;
    if (rc != 0) throw std::runtime_error("close error on '" + path_ + "'");
}

void write_rgbe(const std::string& path, const float* rgba, int width, int height, float scale) {
    if (!rgba || width <= 0 || height <= 0) throw std::runtime_error("empty image");
    RgbeWriter w(path, width, height);
    w.write_rows(rgba, height, scale);
    w.finish();
}

void ExposureSum::add(const float* rgba, long long n) {
    for (long long i = 0; i < n; ++i) {
        const double L = 0.2126 * rgba[4 * i] + 0.7152 * rgba[4 * i + 1] + 0.0722 * rgba[4 * i + 2];
        if (L > 1e-10) {
            log_sum += std::log(1e-6 + L);
            ++lit;
        }
    }
}

float ExposureSum::scale() const {
    const double L_avg = lit > 0 ? std::exp(log_sum / lit) : 1.0;
    return static_cast<float>(0.18 / L_avg);
}

float hdr_exposure_scale(const float* rgba, int width, int height) {
    ExposureSum s;
    s.add(rgba, static_cast<long long>(width) * height);
    return s.scale();
}

}  // namespace grrt_amd

namespace grrt_amd {

namespace {

[[noreturn]] void bad_texture(const std::string& path, const std::string& why) {
    throw std::runtime_error("background texture " + path + ": " + why);
}

// One header line (without the newline, or a CR LF pair); false at end of
// file.
bool read_line(std::FILE* f, std::string& line) {
    line.clear();
    int c;
    while ((c = std::fgetc(f)) != EOF) {
        if (c == '\n') {
            if (!line.empty() && line.back() == '\r') line.pop_back();
            return true;
        }
        line.push_back(static_cast<char>(c));
        if (line.size() > 4096) return false;
    }
    return !line.empty();
}

Texture read_rgbe(std::FILE* f, const std::string& path) {
    std::string line;
    for (;;) {   // header lines up to the blank line
        if (!read_line(f, line)) bad_texture(path, "truncated Radiance header");
        if (line.empty()) break;
        if (line.rfind("FORMAT=", 0) == 0 && line != "FORMAT=32-bit_rle_rgbe")
            bad_texture(path, "unsupported " + line);
    }
    if (!read_line(f, line)) bad_texture(path, "missing resolution line");
    int W = 0, H = 0;
    char ys = 0, xs = 0;
    if (std::sscanf(line.c_str(), "%cY %d %cX %d", &ys, &H, &xs, &W) != 4 || ys != '-' || xs != '+' || W <= 0 ||
        H <= 0 || W > (1 << 16) || H > (1 << 16))
        bad_texture(path, "resolution line must be -Y height +X width: " + line);
    Texture t;
    t.width = W;
    t.height = H;
    t.rgb.assign(static_cast<size_t>(W) * H * 3, 0.0f);
    std::vector<unsigned char> px(static_cast<size_t>(W) * 4);
    for (int y = 0; y < H; ++y) {
        unsigned char head[4];
        if (std::fread(head, 1, 4, f) != 4) bad_texture(path, "truncated pixel data");
        const bool rle = W >= 8 && W < 32768 && head[0] == 2 && head[1] == 2 && ((head[2] << 8) | head[3]) == W &&
                         !(head[2] & 0x80);
        if (rle) {
            for (int ch = 0; ch < 4; ++ch) {
                int x = 0;
                while (x < W) {
                    int count = std::fgetc(f);
                    if (count == EOF) bad_texture(path, "truncated run-length scanline");
                    if (count > 128) {
                        count -= 128;
                        const int v = std::fgetc(f);
                        if (v == EOF || x + count > W) bad_texture(path, "bad run in scanline");
                        for (int k = 0; k < count; ++k) px[static_cast<size_t>(x++) * 4 + ch] = static_cast<unsigned char>(v);
                    } else {
                        if (count == 0 || x + count > W) bad_texture(path, "bad literal in scanline");
                        for (int k = 0; k < count; ++k) {
                            const int v = std::fgetc(f);
                            if (v == EOF) bad_texture(path, "truncated scanline");
                            px[static_cast<size_t>(x++) * 4 + ch] = static_cast<unsigned char>(v);
                        }
                    }
                }
            }
        } else {   // flat scanline: the four bytes read are its first pixel
            std::copy(head, head + 4, px.begin());
            if (W > 1 && std::fread(px.data() + 4, 1, static_cast<size_t>(W - 1) * 4, f) != static_cast<size_t>(W - 1) * 4)
                bad_texture(path, "truncated flat scanline");
        }
        float* row = t.rgb.data() + static_cast<size_t>(y) * W * 3;
        for (int x = 0; x < W; ++x) {
            const unsigned char* p = &px[static_cast<size_t>(x) * 4];
            const float m = p[3] ? std::ldexp(1.0f, static_cast<int>(p[3]) - 136) : 0.0f;
            for (int k = 0; k < 3; ++k) row[3 * x + k] = p[3] ? (p[k] + 0.5f) * m : 0.0f;   // Radiance's colr_color
        }
    }
    return t;
}

Texture read_pfm(std::FILE* f, const std::string& path, bool colour) {
    std::string line;
    int W = 0, H = 0;
    double scale = 0.0;
    if (!read_line(f, line) || std::sscanf(line.c_str(), "%d %d", &W, &H) != 2 || W <= 0 || H <= 0 ||
        W > (1 << 16) || H > (1 << 16))
        bad_texture(path, "bad PFM size line");
    if (!read_line(f, line) || std::sscanf(line.c_str(), "%lf", &scale) != 1 || scale == 0.0)
        bad_texture(path, "bad PFM scale line");
    const bool little = scale < 0.0;
    const int nc = colour ? 3 : 1;
    std::vector<unsigned char> raw(static_cast<size_t>(W) * H * nc * 4);
    if (std::fread(raw.data(), 1, raw.size(), f) != raw.size()) bad_texture(path, "truncated PFM data");
    Texture t;
    t.width = W;
    t.height = H;
    t.rgb.assign(static_cast<size_t>(W) * H * 3, 0.0f);
    for (int y = 0; y < H; ++y)          // stored bottom row first
        for (int x = 0; x < W; ++x)
            for (int k = 0; k < 3; ++k) {
                const unsigned char* b = &raw[((static_cast<size_t>(H - 1 - y) * W + x) * nc + (colour ? k : 0)) * 4];
                const uint32_t u = little ? (uint32_t(b[0]) | uint32_t(b[1]) << 8 | uint32_t(b[2]) << 16 | uint32_t(b[3]) << 24)
                                          : (uint32_t(b[3]) | uint32_t(b[2]) << 8 | uint32_t(b[1]) << 16 | uint32_t(b[0]) << 24);
                float v;
                std::memcpy(&v, &u, 4);
                t.rgb[(static_cast<size_t>(y) * W + x) * 3 + k] = v;
            }
    return t;
}

}  // namespace

Texture read_texture(const std::string& path) {
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) bad_texture(path, "cannot open");
    Texture t;
    try {
        std::string first;
        if (!read_line(f, first)) bad_texture(path, "empty file");
        if (first == "#?RADIANCE" || first == "#?RGBE") t = read_rgbe(f, path);
        else if (first == "PF" || first == "Pf") t = read_pfm(f, path, first == "PF");
        else bad_texture(path, "not a Radiance .hdr or PFM image");
    } catch (...) {
        std::fclose(f);
        throw;
    }
    std::fclose(f);
    for (float& v : t.rgb)
        if (!(v >= 0.0f) || !std::isfinite(v)) v = 0.0f;   // a background radiance is finite and non-negative
    return t;
}

}  // namespace grrt_amd
