"""Output side: file names (src/swspectrometer.cpp:234-254, 487-514) and the
FileSink formats (src/FileSink.cpp:30-107), on the CPU."""
import numpy as np
import pytest

from swspec_amd import sinks


def test_cfg_to_filename_literals():
    # spaces -> '_', every occurrence replaced, integration time truncated to int
    p = "Proj A %fftpoints%_%integrtime%_%channel%_%fftpoints%"
    assert sinks.cfg_to_filename(p, 32000, 1.6, 1) == "Proj_A_32000_1_2_32000"
    assert sinks.cfg_to_filename("plain", 4096, 10.0, 0) == "plain"
    assert sinks.cfg_to_filename(sinks.DEFAULT_PATTERN, 3200000, 10.0, 3) == \
        "ProjDate_StationID_Instrument_ScanNo_3200000_10_4"


def test_output_names_base2_from_pattern1():
    class S:   # the fields output_names reads
        keys = {"BaseFilename1": "obs_%channel%", "FFTIntegrationTimeSec": 2.0}
        fft_points = 1024
        use_channel_file1 = 0
        use_channel_file2 = 2
    n = sinks.output_names(S())
    # base 2 comes from the BaseFilename1 pattern with file 2's channel (:235)
    assert n == {"swspec1": "obs_1_swspec.bin", "swspec2": "obs_3_swspec.bin", "pcal1": "obs_1_pcal.bin",
                 "pcal2": "obs_3_pcal.bin", "xpol": "obs_1_xpol_swspec.bin"}


class _Buf:   # a Buffer stand-in: getData / getLength
    def __init__(self, a, length):
        self.a = np.asarray(a)
        self.n = length

    def getData(self):
        return self.a

    def getLength(self):
        return self.n


def test_binary_sink_appends_length_bytes(tmp_path):
    f = tmp_path / "x_swspec.bin"
    k = sinks.FileSink(sinks.BINARY)
    assert k.write(_Buf(np.ones(4, np.float32), 16)) == 0   # not open: nothing taken
    k.open(str(f))
    s1 = np.arange(5, dtype=np.float32)
    s2 = -np.arange(5, dtype=np.float32)
    assert k.write(_Buf(np.concatenate([s1, [7, 7]]).astype(np.float32), 20)) == 20   # length, not allocation
    assert k.write(_Buf(s2, 20)) == 20
    assert k.write(_Buf(s2, 0)) == 0                                                    # empty buffer
    k.close()
    d = np.fromfile(f, dtype="<f4")
    assert np.array_equal(d, np.concatenate([s1, s2]))
    # the reference's reader (plot_swfile.m) view: spectra of N/2+1 floats
    assert np.array_equal(sinks.read_spectra(str(f), 8), np.stack([s1, s2]))
    # open truncates
    k.open(str(f))
    k.close()
    assert f.stat().st_size == 0


def test_ascii_sink_known_text(tmp_path):
    f = tmp_path / "a.txt"
    k = sinks.FileSink(sinks.ASCII)
    k.open(str(f))
    a = np.array([1.5, 2.25, 3.0, 0.1, 1e-7, 123456.7], dtype=np.float32)
    k.write(_Buf(a, a.nbytes))
    k.close()
    # 24 bytes = 3 complex pairs: floats 0, 2, 4 with 12 significant digits,
    # then float 1 as the Nyquist value (FileSink.cpp:72-90)
    assert f.read_text() == ("// --------------------- DATA SET TIMESTAMP xxxx ---------------------\n"
                             "// FFT bins 0..2 + Nyquist\n"
                             "1.5\n3\n1.00000001169e-07\n2.25\n")


def test_ascii_format_matches_iostream_precision_12():
    # ostream << float with precision(12): the float is promoted to double
    # and printed like %.12g
    assert sinks._fmt12(np.float32(0.1)) == "0.10000000149"
    assert sinks._fmt12(np.float32(123456.7)) == "123456.703125"
    assert sinks._fmt12(np.float32(3.3359)) == "3.33590006828"


# ---------------------------------------------------------------- vs the reference's own FileSink

def _ref_available():
    from oracle import oracle as O
    try:
        return O.build_ref()
    except Exception:
        return False


def _write_both(tmp_path, fmt, arrays, lengths):
    from oracle import oracle as O
    ref, ours = tmp_path / ("ref_" + fmt), tmp_path / ("ours_" + fmt)
    got_ref = O.ref_filesink(fmt, arrays, ref, lengths)
    k = sinks.FileSink(fmt)
    k.open(str(ours))
    got = [k.write(_Buf(np.asarray(a, np.float32), n)) for a, n in zip(arrays, lengths)]
    k.close()
    return ref.read_bytes(), ours.read_bytes(), got_ref, got


@pytest.mark.parametrize("fmt", ["Binary", "Ascii"])
def test_sink_bytes_match_reference_filesink(tmp_path, fmt):
    """sinks.FileSink against src/FileSink.cpp itself (built into oracle/_ref
    from the reference sources with g++): random spectra of odd and even
    lengths written in turn, an empty buffer and a length shorter than the
    array (getLength, not the allocation) -- identical files and returns."""
    if not _ref_available():
        pytest.skip("reference sources not present: oracle/_ref not buildable")
    rng = np.random.default_rng(11)
    arrays = [rng.standard_normal(17).astype(np.float32) * 1e3,
              (rng.random(16001) * 1e-3).astype(np.float32),
              np.zeros(8, np.float32),
              rng.standard_normal(10).astype(np.float32)]
    lengths = [arrays[0].nbytes, arrays[1].nbytes, 0, 24]
    ref, ours, rr, rg = _write_both(tmp_path, fmt, arrays, lengths)
    assert rr == rg == [lengths[0], lengths[1], 0, 24]
    assert ours == ref


def test_ascii_special_values_match_reference_filesink(tmp_path):
    """Values iostream prints in its own ways: -0, subnormals, powers of ten
    at the 12-digit boundary, the largest float, +-inf and NaN."""
    if not _ref_available():
        pytest.skip("reference sources not present: oracle/_ref not buildable")
    a = np.array([-0.0, 1e-45, 1e-38, 1e12, 123456789012.0, 3.4028235e38, np.inf, -np.inf, np.nan, 0.5,
                  1e-5, 100.0], dtype=np.float32)
    ref, ours, rr, rg = _write_both(tmp_path, "Ascii", [a], [a.nbytes])
    assert rr == rg
    assert ours.decode() == ref.decode()
