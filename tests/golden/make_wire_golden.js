// Step 2 of the wire-format fixture (Node v12, build container only):
//   node tests/golden/make_wire_golden.js /root/reference/web/js/net.js
// Decodes tests/golden/wire_frames.bin with the reference viewer's decodeFrame
// and encodes sample commands with its frameBytes; writes wire_golden.json
// (decoded headers and parts, command bytes as hex). The reference file is read
// and evaluated in memory; nothing of it is written.
'use strict';
const fs = require('fs');
const path = require('path');
const src = fs.readFileSync(process.argv[2] || '/root/reference/web/js/net.js', 'utf8');
// evaluate only the two functions used here (the file's Remote class needs a newer Node)
function extract(name) {
  const start = src.indexOf('function ' + name + '(');
  let i = src.indexOf('{', start), depth = 0;
  for (; i < src.length; i++) {
    if (src[i] === '{') depth++;
    else if (src[i] === '}' && --depth === 0) break;
  }
  return src.slice(start, i + 1);
}
const mod = new Function(extract('frameBytes') + '\n' + extract('decodeFrame') + '\nreturn { decodeFrame, frameBytes };')();
const dir = __dirname;
const index = JSON.parse(fs.readFileSync(path.join(dir, 'wire_frames.json'), 'utf8'));
const blob = fs.readFileSync(path.join(dir, 'wire_frames.bin'));
const frames = index.map((c) => {
  const ab = new ArrayBuffer(c.bytes);
  new Uint8Array(ab).set(blob.subarray(c.offset, c.offset + c.bytes));
  const d = mod.decodeFrame(ab);
  const out = { name: c.name, header: d.header };
  for (const k of ['parts16', 'parts8', 'slice', 'cp']) if (d[k]) out[k] = Array.from(d[k]);
  return out;
});
const cmds = [
  [{ type: 'solid', dims: [8, 4, 6] }, Uint8Array.from([1, 2, 3, 4, 5, 250])],
  [{ type: 'surface', numVerts: 2 }, new Uint8Array(new Float32Array([1, 2, 3, 4, 5, 6, 0, 0, 1, 0, 1, 0]).buffer)],
].map(([h, p]) => ({ header: h, payload: Array.from(p), bytes: Buffer.from(mod.frameBytes(h, p)).toString('hex') }));
fs.writeFileSync(path.join(dir, 'wire_golden.json'), JSON.stringify({ generator: 'tests/golden/make_wire_golden.js',
  reference: 'web/js/net.js', node: process.version, frames, commands: cmds }, null, 1) + '\n');
console.log(`decoded ${frames.length} frames, encoded ${cmds.length} commands`);
