"""Output side of the spectrometer: file names and the file sink formats.

- cfg_to_filename: the base-name pattern substitution of
  src/swspectrometer.cpp:487-514 (spaces -> '_', %fftpoints%, %integrtime%
  as an integer, %channel% as the 1-based channel).
- output_names: the files main() opens (src/swspectrometer.cpp:234-254):
  <base1>_swspec.bin, <base2>_swspec.bin, <base1>_pcal.bin,
  <base2>_pcal.bin, <base1>_xpol_swspec.bin.  As in the reference, base 2 is
  built from the BaseFilename1 pattern with file 2's channel (:235).
- FileSink: src/FileSink.cpp:30-107.  Binary appends the Buffer's bytes
  (getLength of them): float32 spectra of fft_ssb_points values, cross-power
  complex64 pairs, phase-cal complex64 tone bins, little-endian.  Ascii writes
  a data-set header line, a bin-range line, then every other float of the
  buffer (its length counted in complex pairs) with 12 significant digits and
  finally the buffer's second float as the Nyquist value, as the reference
  does for a Perm-layout buffer.

Host-side only: the sinks take the host Buffers a core's join() filled (or
numpy arrays), exactly what TaskDispatcher hands the reference's sinks
(src/TaskDispatcher.cpp:142-191).
"""
import numpy as np

BINARY, ASCII = "Binary", "Ascii"


def cfg_to_filename(pattern, fft_points, fft_integ_seconds, channel0):
    """src/swspectrometer.cpp:487-514; channel0 is the 0-based channel."""
    name = pattern.replace(" ", "_")
    keys = (("%fftpoints%", str(int(fft_points))),
            ("%integrtime%", str(int(fft_integ_seconds))),
            ("%channel%", str(int(channel0) + 1)))
    for match, rep in keys:
        while match in name:
            name = name.replace(match, rep, 1)
    return name


DEFAULT_PATTERN = "ProjDate_StationID_Instrument_ScanNo_%fftpoints%_%integrtime%_%channel%"


def output_names(settings, pattern=None):
    """Files main() opens for a derived Settings (src/swspectrometer.cpp:234-254)."""
    pattern = settings.keys.get("BaseFilename1", DEFAULT_PATTERN) if pattern is None else pattern
    T = settings.keys["FFTIntegrationTimeSec"]
    base1 = cfg_to_filename(pattern, settings.fft_points, T, settings.use_channel_file1)
    base2 = cfg_to_filename(pattern, settings.fft_points, T, settings.use_channel_file2)
    return {"swspec1": base1 + "_swspec.bin", "swspec2": base2 + "_swspec.bin",
            "pcal1": base1 + "_pcal.bin", "pcal2": base2 + "_pcal.bin",
            "xpol": base1 + "_xpol_swspec.bin"}


class FileSink:
    """src/FileSink.cpp:30-107: open truncates, write appends one Buffer."""

    def __init__(self, fmt=BINARY):
        if fmt not in (BINARY, ASCII):
            raise ValueError("sink format must be Binary or Ascii")
        self.fmt = fmt
        self.f = None

    def open(self, uri):
        self.close()
        self.f = open(uri, "wb")
        return 0

    def write(self, buf, length=None):
        """buf: a Buffer (its getLength() bytes are written) or a numpy array
        (all of it, or `length` bytes).  Returns the bytes taken, 0 for an
        empty buffer or a closed sink (FileSink.cpp:55-63)."""
        if hasattr(buf, "getData"):
            raw = np.asarray(buf.getData()).view(np.uint8)[: buf.getLength()]
        else:
            raw = np.ascontiguousarray(buf).view(np.uint8).reshape(-1)
            if length is not None:
                raw = raw[:length]
        if raw.size == 0 or self.f is None:
            return 0
        if self.fmt == BINARY:
            self.f.write(raw.tobytes())
        else:
            src = raw[: (raw.size // 4) * 4].view("<f4")
            n = raw.size // 8   # getLength / (2 sizeof(float))
            lines = ["// --------------------- DATA SET TIMESTAMP xxxx ---------------------",
                     "// FFT bins 0..%d + Nyquist" % ((n - 1) % (1 << 64))]   # size_t
            lines += [_fmt12(src[2 * i]) for i in range(n)]
            lines.append(_fmt12(src[1]) if src.size > 1 else _fmt12(0.0))
            self.f.write(("\n".join(lines) + "\n").encode())
        self.f.flush()
        return int(raw.size)

    def close(self):
        if self.f is not None:
            self.f.close()
            self.f = None
        return 0


def _fmt12(x):
    """std::ostream << float with precision(12), default floatfield: %g with
    12 significant digits of the float's (double-promoted) value."""
    return "%.12g" % float(np.float32(x))


def read_spectra(path, fft_points):
    """src/matlab/plot_swfile.m:19-29: float32 little-endian spectra of
    floor(N/2)+1 points each, back to back -> [nspectra, N/2+1] array."""
    ssb = fft_points // 2 + 1
    d = np.fromfile(path, dtype="<f4")
    return d[: (d.size // ssb) * ssb].reshape(-1, ssb)
