"""Run provenance: the reproducibility manifest and the chained-SHA-256 replay log.

Restates the on-disk formats of the reference's ``route34/repro.rs`` so that a run of this
library is audited the same way as a run of the reference:

* :class:`ReproManifest` (``repro.rs:48-71``, ``collect`` at ``repro.rs:156-173``): build and host
  facts captured at run start.
* :class:`ReplogWriter` (``repro.rs:225-345``): a stream of structured events, each paired with
  ``chain[i] = sha256(chain[i-1] || event_json[i])`` where ``chain[-1] = sha256(b"")`` and the
  chain values are the raw 32-byte digests. ``write_to_path`` writes one JSON record per line,
  ``{"seq", "chain_sha256_hex", "event"}``, then the trailer
  ``{"final_chain_sha256_hex", "n_events", "trailer": true}``, and a binary
  ``<stem>.kernel.replog`` sidecar with one fixed-layout entry per event.

The hashed payload of an event is its compact JSON encoding, byte for byte as the reference's
``serde_json::to_vec`` writes it; :func:`to_json_bytes` produces those bytes:

* struct-like objects keep their field order (the dataclasses below hold the reference's order),
  while free-form JSON values (``config_json``, ``summary``) are written with their object keys
  sorted (serde_json's default ``Map`` is ordered by key);
* a float is written in the shortest form that reads back to the same double, laid out as the
  Ryu formatter lays it out (``1e-6``, ``0.0000613``, ``1234.5``, ``1.0``, ``1.5e300``), which
  differs from Python's ``repr`` for small and large magnitudes; non-finite floats are ``null``;
* strings escape only ``"``, ``\\`` and control characters.

The format of the sidecar (``PWOSRPL1``) is that of the files the reference writes: an 8-byte
magic, ``u32`` version 1, ``u32`` entry count, then per entry ``u32`` name length, the name,
``u64`` seed (the event's sequence number), 32-byte input SHA-256 (of the event JSON), 32-byte
output SHA-256 (the chain value), ``u32`` event-JSON length, ``u32`` 32, ``u64`` 0; all
little-endian.
"""
from __future__ import annotations

import hashlib
import json
import math
import os
import platform
import struct
import subprocess
import sys
import time
from dataclasses import dataclass, field, fields

__all__ = [
    "format_f64", "to_json_bytes", "sorted_value", "sha256", "rfc3339_utc_now",
    "ReproManifest", "PerSeedEvent", "run_start_event", "per_seed_event", "run_end_event",
    "ReplogWriter", "ReplogError", "read_replog", "verify_replog", "kernel_replog_bytes",
    "parse_kernel_replog", "KERNEL_REPLOG_MAGIC",
]

KERNEL_REPLOG_MAGIC = b"PWOSRPL1"
_KERNEL_EVENT_NAME = b"discriminator_event"


class ReplogError(ValueError):
    """A replay log that does not verify (broken chain, bad trailer, malformed line)."""


# ---- canonical JSON ---------------------------------------------------------------------------

def format_f64(x: float) -> str:
    """The JSON text serde_json writes for the double ``x``.

    The digits are the shortest decimal string that reads back as ``x`` (Python's ``repr`` finds
    the same digit string). They are laid out by the Ryu rule on ``kk``, the decimal exponent of
    the leading digit plus one: integers below 1e16 as ``digits.0``, ``0 < kk <= 16`` as a plain
    decimal, ``-5 < kk <= 0`` as ``0.000digits``, otherwise ``d.ddde<kk-1>`` with no ``+`` and no
    zero padding in the exponent."""
    if not math.isfinite(x):
        return "null"
    if x == 0.0:
        return "-0.0" if math.copysign(1.0, x) < 0 else "0.0"
    sign = "-" if x < 0 else ""
    mant, _, exp = repr(abs(x)).partition("e")
    e10 = int(exp) if exp else 0
    ip, _, fp = mant.partition(".")
    digits = (ip + fp).lstrip("0")
    e10 -= len(fp)
    stripped = digits.rstrip("0")
    e10 += len(digits) - len(stripped)
    digits = stripped
    n = len(digits)
    kk = n + e10
    if 0 <= e10 and kk <= 16:
        body = digits + "0" * e10 + ".0"
    elif 0 < kk <= 16:
        body = digits[:kk] + "." + digits[kk:]
    elif -5 < kk <= 0:
        body = "0." + "0" * (-kk) + digits
    elif n == 1:
        body = digits + "e" + str(kk - 1)
    else:
        body = digits[0] + "." + digits[1:] + "e" + str(kk - 1)
    return sign + body


def _json_str(s: str) -> str:
    return json.dumps(s, ensure_ascii=False)


def _encode(v, out: list):
    if v is None:
        out.append("null")
    elif v is True:
        out.append("true")
    elif v is False:
        out.append("false")
    elif isinstance(v, int):
        out.append(str(int(v)))
    elif isinstance(v, float):
        out.append(format_f64(v))
    elif isinstance(v, str):
        out.append(_json_str(v))
    elif isinstance(v, dict):
        out.append("{")
        first = True
        for k, x in v.items():
            if not isinstance(k, str):
                raise TypeError(f"JSON object keys must be str, got {type(k).__name__}")
            if not first:
                out.append(",")
            first = False
            out.append(_json_str(k))
            out.append(":")
            _encode(x, out)
        out.append("}")
    elif isinstance(v, (list, tuple)):
        out.append("[")
        for i, x in enumerate(v):
            if i:
                out.append(",")
            _encode(x, out)
        out.append("]")
    elif hasattr(v, "tolist"):  # numpy scalars and arrays
        _encode(v.tolist(), out)
    else:
        raise TypeError(f"cannot encode {type(v).__name__} as JSON")


def to_json_bytes(value) -> bytes:
    """Compact JSON (no whitespace) of ``value``; dict keys are written in insertion order."""
    out: list = []
    _encode(value, out)
    return "".join(out).encode("utf-8")


def sorted_value(v):
    """``v`` with every object's keys in sorted order: the order in which serde_json writes a
    free-form ``serde_json::Value``."""
    if isinstance(v, dict):
        return {k: sorted_value(v[k]) for k in sorted(v)}
    if isinstance(v, (list, tuple)):
        return [sorted_value(x) for x in v]
    if hasattr(v, "tolist") and not isinstance(v, (str, bytes)):
        return sorted_value(v.tolist())
    return v


def sha256(data: bytes) -> bytes:
    return hashlib.sha256(data).digest()


def rfc3339_utc_now(now: float | None = None) -> str:
    """``YYYY-MM-DDTHH:MM:SSZ`` in UTC, whole seconds (repro.rs:121-146)."""
    t = time.gmtime(int(time.time() if now is None else now))
    return time.strftime("%Y-%m-%dT%H:%M:%SZ", t)


# ---- manifest ---------------------------------------------------------------------------------

_X86_FEATURES = [  # (name written, /proc/cpuinfo flag), in the reference's order (repro.rs:92-95)
    ("sse2", "sse2"), ("sse4.1", "sse4_1"), ("sse4.2", "sse4_2"), ("avx", "avx"),
    ("avx2", "avx2"), ("fma", "fma"), ("avx512f", "avx512f"), ("avx512vl", "avx512vl"),
    ("avx512bw", "avx512bw"), ("avx512dq", "avx512dq"),
]


def _cpu_features() -> list:
    if platform.machine() not in ("x86_64", "AMD64"):
        return []
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    flags = set(line.split(":", 1)[1].split())
                    return [name for name, flag in _X86_FEATURES if flag in flags]
    except OSError:
        pass
    return []


def _git_revision(cwd: str | None) -> str | None:
    try:
        r = subprocess.run(["git", "rev-parse", "HEAD"], capture_output=True, text=True,
                           cwd=cwd, timeout=10)
    except (OSError, subprocess.SubprocessError):
        return None
    return r.stdout.strip() if r.returncode == 0 else None


def _target_triple() -> str:
    m = platform.machine() or "unknown"
    s = sys.platform
    if s.startswith("linux"):
        return f"{m}-unknown-linux-gnu"
    if s == "win32":
        return f"{m}-pc-windows-msvc"
    if s == "darwin":
        return f"{m}-apple-darwin"
    return f"{m}-unknown-{s}"


def _library_toolchain() -> str:
    """The version string of the loaded compute library, which names the HIP compiler that built
    it and the GPU target; "unknown" when the library cannot be loaded in this process."""
    try:
        from . import quintic
        return quintic.library_version()
    except Exception:  # the manifest is metadata: a missing library is recorded, not fatal
        return "unknown"


@dataclass
class ReproManifest:
    """ReproManifest (repro.rs:48-71). The first eight fields keep the reference's names and
    order so that either side's reader accepts the other's files; ``rust_toolchain`` holds the
    toolchain that built the compute library (for this library: the hipcc/clang version) and
    ``n_threads_rayon`` the host threads this process may use. The fields after them are this
    library's additions (readers ignore unknown fields)."""
    git_revision: str | None
    rust_toolchain: str
    target_triple: str
    cpu_features: list
    hostname: str | None
    timestamp_utc: str
    command_line: list
    n_threads_rayon: int
    gpu_arch: str = "gfx950"
    python_version: str = ""

    @classmethod
    def collect(cls, repo_dir: str | None = None) -> "ReproManifest":
        """ReproManifest::collect (repro.rs:156-173)."""
        try:
            n_threads = len(os.sched_getaffinity(0))
        except AttributeError:
            n_threads = os.cpu_count() or 1
        return cls(
            git_revision=_git_revision(repo_dir or os.path.dirname(os.path.abspath(__file__))),
            rust_toolchain=_library_toolchain(),
            target_triple=_target_triple(),
            cpu_features=_cpu_features(),
            hostname=os.environ.get("COMPUTERNAME") or os.environ.get("HOSTNAME"),
            timestamp_utc=rfc3339_utc_now(),
            command_line=list(sys.argv),
            n_threads_rayon=max(1, n_threads),
            python_version=platform.python_version(),
        )

    def to_dict(self) -> dict:
        d = {}
        for f in fields(self):
            d[f.name] = getattr(self, f.name)
        return d

    @classmethod
    def from_dict(cls, d: dict) -> "ReproManifest":
        """Reads a manifest written by either side: unknown fields are dropped, this library's
        additions are left out of the object when the file has none (so that re-encoding the
        object reproduces the file)."""
        known = {f.name for f in fields(cls)}
        m = cls(**{k: v for k, v in d.items() if k in known})
        m._present = [k for k in d if k in known]
        return m


def _manifest_dict(m) -> dict:
    if isinstance(m, ReproManifest):
        d = m.to_dict()
        present = getattr(m, "_present", None)
        return {k: d[k] for k in present} if present is not None else d
    return dict(m)


@dataclass
class PerSeedEvent:
    """PerSeedEvent (repro.rs:199-210): one metric solve."""
    seed: int
    candidate: str
    k: int
    iters_run: int
    final_residual: float
    sigma_fs_identity: float
    sigma_final: float
    n_basis: int
    elapsed_ms: float

    def to_dict(self) -> dict:
        out = {}
        for f in fields(self):
            v = getattr(self, f.name)
            out[f.name] = float(v) if f.type == "float" else v
        return out


# ReplogEvent (repro.rs:180-196): internally tagged by "kind", variant names in snake_case.

def run_start_event(binary: str, manifest, config_json) -> dict:
    return {"kind": "run_start", "binary": str(binary), "manifest": _manifest_dict(manifest),
            "config_json": sorted_value(config_json)}


def per_seed_event(ev: PerSeedEvent) -> dict:
    return {"kind": "per_seed", **ev.to_dict()}


def run_end_event(summary, total_elapsed_s: float) -> dict:
    return {"kind": "run_end", "summary": sorted_value(summary),
            "total_elapsed_s": float(total_elapsed_s)}


# ---- writer -----------------------------------------------------------------------------------

class ReplogWriter:
    """ReplogWriter (repro.rs:225-345): appends events and advances the SHA-256 chain."""

    def __init__(self):
        self.records: list = []  # (seq, chain digest, event bytes, event)
        self._chain = sha256(b"")
        self._seq = 0

    def push(self, event: dict) -> str:
        """Hash the event's canonical JSON into the chain; returns the new chain value (hex)."""
        if not isinstance(event, dict) or event.get("kind") not in ("run_start", "per_seed",
                                                                     "run_end"):
            raise ValueError("event must be a run_start, per_seed or run_end event dict")
        eb = to_json_bytes(event)
        self._chain = sha256(self._chain + eb)
        self.records.append((self._seq, self._chain, eb, event))
        self._seq += 1
        return self._chain.hex()

    def final_chain_hex(self) -> str:
        return self._chain.hex()

    def to_lines(self) -> bytes:
        """The JSON Lines body: one record per event, then the trailer."""
        out = []
        for seq, chain, eb, _ in self.records:
            out.append(b'{"seq":%d,"chain_sha256_hex":"%s","event":' % (seq, chain.hex().encode())
                       + eb + b"}\n")
        trailer = {"final_chain_sha256_hex": self._chain.hex(), "n_events": len(self.records),
                   "trailer": True}
        out.append(to_json_bytes(trailer) + b"\n")
        return b"".join(out)

    def kernel_sidecar(self) -> bytes:
        return kernel_replog_bytes([(seq, sha256(eb), chain, len(eb))
                                    for seq, chain, eb, _ in self.records])

    def write_to_path(self, path: str) -> str:
        """ReplogWriter::write_to_path (repro.rs:306-345): the JSON Lines file at ``path`` and
        the binary sidecar beside it (``<stem>.kernel.replog``). Both are written to a temporary
        name and renamed into place. Returns the sidecar path."""
        parent = os.path.dirname(os.path.abspath(path))
        os.makedirs(parent, exist_ok=True)
        _atomic_write(path, self.to_lines())
        side = os.path.splitext(path)[0] + ".kernel.replog"
        _atomic_write(side, self.kernel_sidecar())
        return side


def _atomic_write(path: str, data: bytes):
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(data)
    os.replace(tmp, path)


def kernel_replog_bytes(entries) -> bytes:
    """The ``PWOSRPL1`` sidecar for ``(seq, input_sha, output_sha, n_event_bytes)`` entries."""
    out = [KERNEL_REPLOG_MAGIC, struct.pack("<II", 1, len(entries))]
    for seq, in_sha, out_sha, n_in in entries:
        if len(in_sha) != 32 or len(out_sha) != 32:
            raise ValueError("SHA-256 digests are 32 bytes")
        out.append(struct.pack("<I", len(_KERNEL_EVENT_NAME)) + _KERNEL_EVENT_NAME)
        out.append(struct.pack("<Q", seq) + bytes(in_sha) + bytes(out_sha))
        out.append(struct.pack("<IIQ", min(n_in, 0xFFFFFFFF), 32, 0))
    return b"".join(out)


def parse_kernel_replog(data: bytes) -> list:
    """Entries of a ``PWOSRPL1`` sidecar as dicts (name, seed, input_sha, output_sha, n_in,
    n_out, t_ns)."""
    if data[:8] != KERNEL_REPLOG_MAGIC:
        raise ReplogError("not a PWOSRPL1 file")
    ver, n = struct.unpack_from("<II", data, 8)
    if ver != 1:
        raise ReplogError(f"unsupported sidecar version {ver}")
    pos, out = 16, []
    for _ in range(n):
        (ln,) = struct.unpack_from("<I", data, pos)
        pos += 4
        name = data[pos:pos + ln]
        pos += ln
        (seed,) = struct.unpack_from("<Q", data, pos)
        pos += 8
        in_sha, out_sha = data[pos:pos + 32], data[pos + 32:pos + 64]
        pos += 64
        n_in, n_out, t_ns = struct.unpack_from("<IIQ", data, pos)
        pos += 16
        out.append({"name": name.decode(), "seed": seed, "input_sha": in_sha,
                    "output_sha": out_sha, "n_in": n_in, "n_out": n_out, "t_ns": t_ns})
    if pos != len(data):
        raise ReplogError(f"{len(data) - pos} trailing bytes after {n} sidecar entries")
    return out


# ---- reader / verifier ------------------------------------------------------------------------

def read_replog(path: str):
    """(records, trailer) of a replay log: records are (seq, chain hex, raw event bytes, event)."""
    with open(path, "rb") as f:
        lines = [ln for ln in f.read().split(b"\n") if ln]
    records, trailer = [], None
    for i, ln in enumerate(lines):
        rec = json.loads(ln)
        if rec.get("trailer") is True:
            if i != len(lines) - 1:
                raise ReplogError("trailer is not the last line")
            trailer = rec
            continue
        seq, chain = rec.get("seq"), rec.get("chain_sha256_hex")
        prefix = b'{"seq":%d,"chain_sha256_hex":"%s","event":' % (seq, str(chain).encode())
        if not (ln.startswith(prefix) and ln.endswith(b"}")):
            raise ReplogError(f"line {i}: record is not in canonical form")
        records.append((seq, chain, ln[len(prefix):-1], rec["event"]))
    return records, trailer


def verify_replog(path: str, start_chain: bytes | None = None) -> str:
    """Recompute the chain of a replay log and check every record, the trailer and (when present)
    the sidecar. Returns the final chain value (hex); raises :class:`ReplogError` otherwise.

    ``start_chain`` replaces ``sha256(b"")`` as the chain seed; every record must still carry the
    next sequence number."""
    records, trailer = read_replog(path)
    chain = sha256(b"") if start_chain is None else start_chain
    for i, (seq, hexd, eb, ev) in enumerate(records):
        if seq != i:
            raise ReplogError(f"record {i} has seq {seq}")
        chain = sha256(chain + eb)
        if chain.hex() != hexd:
            raise ReplogError(f"chain mismatch at seq {seq}")
    if trailer is None:
        raise ReplogError("missing trailer")
    if trailer.get("n_events") != len(records) or trailer.get(
            "final_chain_sha256_hex") != chain.hex():
        raise ReplogError("trailer does not match the records")
    side = os.path.splitext(path)[0] + ".kernel.replog"
    if os.path.exists(side):
        with open(side, "rb") as f:
            ents = parse_kernel_replog(f.read())
        if len(ents) != len(records):
            raise ReplogError("sidecar entry count differs from the record count")
        for e, (seq, hexd, eb, _) in zip(ents, records):
            if (e["seed"] != seq or e["output_sha"].hex() != hexd
                    or e["input_sha"] != sha256(eb) or e["n_in"] != len(eb)):
                raise ReplogError(f"sidecar entry {seq} does not match its record")
    return chain.hex()
