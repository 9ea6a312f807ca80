'use strict';
// Runs a few molecule.js frames through the JS host on the GPU and writes the
// energies and final fields for the parity test (tests/test_node_addon.py).
const fs = require('fs');
const path = require('path');
const { MoleculeHost } = require(path.join(__dirname, 'molecule_host.js'));

const out = process.argv[2];
const spec = JSON.parse(process.argv[3]);
const host = new MoleculeHost(spec.cfg);
host.initGPU();
if (spec.activate) host.activateAtoms(spec.activate);
const frames = [];
for (let f = 0; f < spec.frames; f++) {
  host.doSteps(spec.steps);
  frames.push({E: host.E, E_T: host.E_T, E_eK: host.E_eK, E_ee: host.E_ee, E_KK: host.E_KK, dipole: host.dipole});
}
const U = host.readField('U');
const L = host.readField('label');
fs.writeFileSync(path.join(out, 'node_U.f32'), Buffer.from(U.buffer));
fs.writeFileSync(path.join(out, 'node_label.i32'), Buffer.from(L.buffer));
fs.writeFileSync(path.join(out, 'node_frames.json'), JSON.stringify(frames));
host.destroy();
console.log('node host: ' + frames.length + ' frames, E=' + frames[frames.length - 1].E.toFixed(8));
