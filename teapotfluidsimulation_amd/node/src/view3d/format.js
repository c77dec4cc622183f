'use strict';
// The pour3d.bin.gz recording of src/view3d/format.ts:1-8, little endian:
//   u32 headerLength | header JSON (UTF-8) | u8 origin[P] | u16 frame0[3P] | i16 delta[F-1][3P]
// The frames arrive already quantised and delta-encoded by the device (Recorder3D.capture), so
// writing a file is concatenation; decodeRecording is the playback side's reader.

// The raw (uncompressed) file from the header, the origin bytes and each frame's Uint16Array.
function writeRecording(header, origin, frameWords) {
  const head = Buffer.from(JSON.stringify(header), 'utf8');
  const P = header.parcels;
  const size = 4 + head.length + P + frameWords.length * 3 * P * 2;
  const buf = Buffer.alloc(size);
  buf.writeUInt32LE(head.length, 0);
  head.copy(buf, 4);
  let o = 4 + head.length;
  Buffer.from(origin.buffer, origin.byteOffset, origin.byteLength).copy(buf, o);
  o += P;
  for (const w of frameWords) {
    if (w.length !== 3 * P) throw new Error(`frame has ${w.length} words, expected ${3 * P}`);
    for (let i = 0; i < w.length; i++, o += 2) buf.writeUInt16LE(w[i], o);
  }
  return buf;
}

// decodeRecording (format.ts:78-99): {header, origin, positions[frame] = Float32Array(3P)}.
function decodeRecording(buf) {
  const dv = new DataView(buf.buffer, buf.byteOffset, buf.byteLength);
  const hl = dv.getUint32(0, true);
  const header = JSON.parse(Buffer.from(buf.buffer, buf.byteOffset + 4, hl).toString('utf8'));
  if (header.version !== 1) throw new Error(`Unsupported recording version ${header.version}`);
  const P = header.parcels;
  let o = 4 + hl;
  const origin = Uint8Array.from(buf.subarray(o, o + P));
  o += P;
  const positions = [];
  const cur = new Int32Array(3 * P);
  for (let k = 0; k < header.frames; k++) {
    const f = new Float32Array(3 * P);
    for (let i = 0; i < 3 * P; i++, o += 2) {
      cur[i] = k === 0 ? dv.getUint16(o, true) : cur[i] + dv.getInt16(o, true);
      f[i] = cur[i] * header.quantum;
    }
    positions.push(f);
  }
  return { header, origin, positions };
}

module.exports = { writeRecording, decodeRecording };
