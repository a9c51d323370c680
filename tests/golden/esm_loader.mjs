// Node 12 module-loader hook used only by gen_golden.mjs: the reference's
// `.js` files are ES modules but ship without a package.json, so Node would
// treat them as CommonJS. This hook reports them as ESM, read in place.
export async function getFormat(url, context, defaultGetFormat) {
  if (url.startsWith('file:///root/reference/') && url.endsWith('.js')) {
    return { format: 'module' };
  }
  return defaultGetFormat(url, context, defaultGetFormat);
}
