"""Trial data files of the eyelid drivers (ECTrialsData), in numpy.

The format and update rules are those of CBM_DATA_LIB_DEV (ectrialsdata.cpp,
peristimhistfloat.cpp, spikerasterbitarray.cpp, ispikeraster.cpp,
eyelidout.cpp, rawuintdata.cpp); the C++ drop-in of the same classes is
csrc/core/ectrialsdata.cpp.  Each object writes a text line of unsigned
integers and then its raw array; ECTrialsData writes its line, the eyelid
data, then the histograms, rasters and raw uint rows, each group preceded by a
count line and each object by a line with its name, in name order.

`TrialDataRecorder` fills a file the way ECManagementDelay::calcSimActivity
does (Eyelid_Sim_CBM_Dev/src/ectrial/ecmanagementdelay.cpp:233-271) -- the
histograms and the eyelid every step of the window [csOn-500, csOff+500),
a raster word every 32 steps -- but from whole trials of spikes taken from
the device recorder instead of one host read per step.
"""
from __future__ import annotations

import math

import numpy as np

f32 = np.float32


def _ints(tokens, n):
    return [int(tokens.pop(0)) for _ in range(n)]


class _Reader:
    """Stream reading with the semantics of std::istream >> and seekg(1, cur)."""

    def __init__(self, data: bytes):
        self.d, self.p = data, 0

    def token(self) -> str:
        while self.d[self.p:self.p + 1].isspace():
            self.p += 1
        q = self.p
        while q < len(self.d) and not self.d[q:q + 1].isspace():
            q += 1
        t = self.d[self.p:q].decode()
        self.p = q
        return t

    def uint(self) -> int:
        return int(self.token())

    def skip1(self):
        self.p += 1

    def array(self, dtype, n):
        b = np.frombuffer(self.d, dtype=dtype, count=n, offset=self.p).copy()
        self.p += b.nbytes
        return b


class PeriStimHistFloat:
    def __init__(self, num_steps: int, steps_per_bin: int, num_cells: int):
        self.nCells, self.numTS, self.tsPerBin = num_cells, num_steps, steps_per_bin
        self.numBins = math.ceil(num_steps / float(f32(steps_per_bin)))
        self.curTS = self.curTrialN = 0
        self.cells = np.zeros((self.numBins, num_cells), f32)

    def update(self, x):
        x = np.asarray(x)
        b = self.cells[self.curTS // self.tsPerBin]
        if x.dtype == np.uint8:
            b += (x[:self.nCells] > 0).astype(f32)
        else:
            b += x[:self.nCells].astype(f32)
        self.curTS += 1
        if self.curTS >= self.numTS:
            self.curTS = 0
            self.curTrialN += 1

    def write(self) -> bytes:
        head = f"{self.nCells} {self.numTS} {self.tsPerBin} {self.numBins} {self.curTS} {self.curTrialN}\n"
        return head.encode() + self.cells.tobytes()

    @classmethod
    def read(cls, r: _Reader):
        self = cls.__new__(cls)
        self.nCells, self.numTS, self.tsPerBin, self.numBins, self.curTS, self.curTrialN = (r.uint() for _ in range(6))
        r.skip1()
        self.cells = r.array(f32, self.numBins * self.nCells).reshape(self.numBins, self.nCells)
        return self


class SpikeRasterBitArray:
    def __init__(self, cell_type: str, num_cells: int, num_trials: int, ms_per_trial: int, ms_per_ts: int):
        self.cellT, self.nCells, self.nTrials, self.msPerTrial = cell_type, num_cells, num_trials, ms_per_trial
        self.msPerTS = ms_per_ts
        self.curTrialN = self.curTS = self.curBufN = 0
        self.tsPerTrial = ms_per_trial // ms_per_ts
        self.bufsPerTrial = self.tsPerTrial // 32
        self.raster = np.zeros((num_trials, self.bufsPerTrial, num_cells), np.uint32)

    def update(self, words) -> bool:
        if self.curTrialN >= self.nTrials:
            return False
        self.raster[self.curTrialN, self.curBufN] = np.asarray(words, np.uint32)[:self.nCells]
        self.curBufN += 1
        self.curTS += 32
        if self.curBufN >= self.bufsPerTrial:
            self.curTrialN += 1
            self.curBufN = self.curTS = 0
        return True

    def spike_times(self, cell: int, trial: int, offset: int = 0) -> list[int]:
        """Bit 31 of a word is the earliest of its 32 steps."""
        out, t = [], offset
        for w in self.raster[trial, :, cell]:
            for j in range(32):
                if (int(w) << j) & 0x80000000:
                    out.append(t)
                t += self.msPerTS
        return out

    def write(self) -> bytes:
        head = (f"{self.cellT} {self.nCells} {self.nTrials} {self.msPerTrial} {self.msPerTS} {self.curTrialN} "
                f"{self.curTS} {self.curBufN} {self.tsPerTrial} {self.bufsPerTrial}\n")
        return head.encode() + self.raster.tobytes()

    @classmethod
    def read(cls, r: _Reader):
        self = cls.__new__(cls)
        self.cellT = r.token()
        self.nCells, self.nTrials, self.msPerTrial = r.uint(), r.uint(), r.uint()
        r.skip1()
        (self.msPerTS, self.curTrialN, self.curTS, self.curBufN, self.tsPerTrial,
         self.bufsPerTrial) = (r.uint() for _ in range(6))
        r.skip1()
        self.raster = r.array(np.uint32, self.nTrials * self.bufsPerTrial * self.nCells).reshape(
            self.nTrials, self.bufsPerTrial, self.nCells)
        return self


class EyelidOut:
    def __init__(self, num_trials: int, ms_per_ts: int, num_ts_smooth: int, ms_per_trial: int):
        self.nTrials, self.msPerTS, self.numTSSmooth, self.msPerTrial = num_trials, ms_per_ts, num_ts_smooth, ms_per_trial
        self.tsPerTrial = ms_per_trial // ms_per_ts
        self.curTrialN = self.curTS = 0
        self.rawData = np.zeros((num_trials, self.tsPerTrial), f32)

    def update(self, x) -> bool:
        """The stored point is the mean of x and the numTSSmooth-1 points stored
        before it in the trial, which are themselves smoothed."""
        if self.curTrialN >= self.nTrials:
            return False
        row = self.rawData[self.curTrialN]
        if self.curTS < self.numTSSmooth - 1:
            row[self.curTS] = f32(x)
        else:
            a = f32(0)
            for j in range(-(self.numTSSmooth - 1), 0):
                a = f32(a + row[self.curTS + j])
            a = f32(a + f32(x))
            row[self.curTS] = f32(a / f32(self.numTSSmooth))
        self.curTS += 1
        if self.curTS >= self.tsPerTrial:
            self.curTrialN += 1
            self.curTS = 0
        return True

    def write(self) -> bytes:
        head = (f"{self.nTrials} {self.msPerTS} {self.numTSSmooth} {self.msPerTrial} {self.tsPerTrial} "
                f"{self.curTrialN} {self.curTS}\n")
        return head.encode() + self.rawData.tobytes()

    @classmethod
    def read(cls, r: _Reader):
        self = cls.__new__(cls)
        (self.nTrials, self.msPerTS, self.numTSSmooth, self.msPerTrial, self.tsPerTrial, self.curTrialN,
         self.curTS) = (r.uint() for _ in range(7))
        r.skip1()
        self.rawData = r.array(f32, self.nTrials * self.tsPerTrial).reshape(self.nTrials, self.tsPerTrial)
        return self


class RawUIntData:
    def __init__(self, label: str, num_rows: int, num_trials: int, points_per_trial: int):
        self.dataLabel, self.numDataRows, self.maxNumTrials, self.numPointsPerTrial = label, num_rows, num_trials, points_per_trial
        self.curTrialN = self.curPoint = 0
        self.data = np.zeros((num_trials, points_per_trial, num_rows), np.uint32)

    def update(self, x) -> bool:
        if self.curTrialN >= self.maxNumTrials:
            return False
        self.data[self.curTrialN, self.curPoint] = np.asarray(x, np.uint32)[:self.numDataRows]
        self.curPoint += 1
        if self.curPoint >= self.numPointsPerTrial:
            self.curTrialN += 1
            self.curPoint = 0
        return True

    def write(self) -> bytes:
        head = (f"{self.dataLabel} {self.numDataRows} {self.maxNumTrials} {self.numPointsPerTrial} "
                f"{self.curTrialN} {self.curPoint}\n")
        return head.encode() + self.data.tobytes()

    @classmethod
    def read(cls, r: _Reader):
        self = cls.__new__(cls)
        self.dataLabel = r.token()
        (self.numDataRows, self.maxNumTrials, self.numPointsPerTrial, self.curTrialN,
         self.curPoint) = (r.uint() for _ in range(5))
        r.skip1()
        self.data = r.array(np.uint32, self.maxNumTrials * self.numPointsPerTrial * self.numDataRows).reshape(
            self.maxNumTrials, self.numPointsPerTrial, self.numDataRows)
        return self


class ECTrialsData:
    """ECTrialsData(msPreCS, msCS, msPostCS, msPerTimeStep, numTrials,
    psh={name: (numTimeStepsPerBin, numCells)}, rasters={name: numCells},
    raw_uint={name: numRows}, eyelid_smooth=numTimeStepSmooth)."""

    def __init__(self, ms_pre: int, ms_cs: int, ms_post: int, ms_per_ts: int, num_trials: int, *, psh=None,
                 rasters=None, raw_uint=None, eyelid_smooth: int = 4):
        self.msPreCS, self.msCS, self.msPostCS, self.msPerTS, self.numTrials = ms_pre, ms_cs, ms_post, ms_per_ts, num_trials
        self.msTotal = ms_pre + ms_cs + ms_post
        self.tsTotal = self.msTotal // ms_per_ts
        self.eyelid = EyelidOut(num_trials, ms_per_ts, eyelid_smooth, self.msTotal)
        self.psh = {k: PeriStimHistFloat(self.tsTotal, b, n) for k, (b, n) in (psh or {}).items()}
        self.rasters = {k: SpikeRasterBitArray(k, n, num_trials, self.msTotal, ms_per_ts)
                        for k, n in (rasters or {}).items()}
        self.uint = {k: RawUIntData(k, n, num_trials, self.tsTotal) for k, n in (raw_uint or {}).items()}

    def write(self) -> bytes:
        out = [f"{self.msPreCS} {self.msCS} {self.msPostCS} {self.msTotal} {self.msPerTS} {self.tsTotal} "
               f"{self.numTrials}\n".encode(), self.eyelid.write()]
        for group in (self.psh, self.rasters, self.uint):
            out.append(f"{len(group)}\n".encode())
            for k in sorted(group):
                out.append(f"{k}\n".encode())
                out.append(group[k].write())
        return b"".join(out)

    @classmethod
    def read(cls, data: bytes) -> "ECTrialsData":
        r = _Reader(data)
        self = cls.__new__(cls)
        (self.msPreCS, self.msCS, self.msPostCS, self.msTotal, self.msPerTS, self.tsTotal,
         self.numTrials) = (r.uint() for _ in range(7))
        r.skip1()
        self.eyelid = EyelidOut.read(r)
        for attr, kind in (("psh", PeriStimHistFloat), ("rasters", SpikeRasterBitArray), ("uint", RawUIntData)):
            group = {}
            for _ in range(r.uint()):
                name = r.token()
                r.skip1()
                group[name] = kind.read(r)
            setattr(self, attr, group)
        return self


def ap_buf_words(spikes: np.ndarray, steps) -> np.ndarray:
    """The spike registers after the given steps (bit k = spike k steps before):
    spikes u8 [T, cells] -> u32 [len(steps), cells], from the 32 steps that end at each."""
    s = (np.asarray(spikes) != 0).astype(np.uint64)
    out = np.zeros((len(steps), s.shape[1]), np.uint64)
    for i, t in enumerate(steps):
        for k in range(32):
            if t - k >= 0:
                out[i] |= s[t - k] << np.uint64(k)
    return out.astype(np.uint32)


class TrialDataRecorder:
    """Fills an ECTrialsData with whole trials, as ECManagementDelay records
    them step by step (ecmanagementdelay.cpp:233-271): for each step ct of the
    window (csOn-500 <= t < csOff+500), the histograms take that step's
    spikes (or values) and the eyelid takes its position; when ct > 0 and
    ct % 32 == 0 the rasters take the spike registers of that step."""

    def __init__(self, data: ECTrialsData):
        self.data = data

    def add_trial(self, *, psh=None, rasters=None, eyelid=None, raw_uint=None):
        """Arrays over the window's steps: psh[name]: [T, cells] (u8 spikes,
        u32 counts or f32 values); rasters[name]: u8 spikes [T, cells];
        eyelid: [T]; raw_uint[name]: u32 [T, rows]."""
        d = self.data
        T = d.tsTotal
        for k, x in (psh or {}).items():
            x = np.asarray(x)
            h = d.psh[k]
            if x.dtype == np.uint8 or np.issubdtype(x.dtype, np.integer):
                # integer steps: exact in FP32 in any order; add the whole trial per bin
                vals = (x[:T, :h.nCells] > 0).astype(f32) if x.dtype == np.uint8 else x[:T, :h.nCells].astype(f32)
                ts = (h.curTS + np.arange(T)) % h.numTS
                np.add.at(h.cells, ts // h.tsPerBin, vals)
                wraps = (h.curTS + T) // h.numTS
                h.curTS = (h.curTS + T) % h.numTS
                h.curTrialN += wraps
            else:
                for t in range(T):
                    h.update(x[t])
        if rasters:
            steps = [ct for ct in range(1, T) if ct % 32 == 0]
            for k, sp in rasters.items():
                words = ap_buf_words(sp[:T], steps)
                for w in words:
                    d.rasters[k].update(w)
        if eyelid is not None:
            for t in range(T):
                d.eyelid.update(eyelid[t])
        for k, x in (raw_uint or {}).items():
            for t in range(T):
                d.uint[k].update(x[t])
