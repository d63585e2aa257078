'use strict';
/** Node.js entry point: the gfx950 solvers and their host-side helpers. */
module.exports = Object.assign({}, require('./solver'), require('./geometry'));
