/* SImage: the image descriptor passed by value across the C ABI.
 *
 * Layout and enum values are identical to the reference descriptor
 * (include/SImage.h:28-48) because callers build it on their side of the ABI:
 * four 32-bit unsigned fields followed by the element type enum.
 */
#ifndef __SIMAGE__H
#define __SIMAGE__H

struct SImage
{
   unsigned int Width;     /* pixels per row */
   unsigned int Height;    /* rows */
   unsigned int Step;      /* bytes from one row to the next (pitch) */
   unsigned int Channels;  /* 1, 2, 3 or 4 interleaved channels */

   enum EDataType
   {
      U8,   /* uint8_t  */
      S8,   /* int8_t   */
      U16,  /* uint16_t */
      S16,  /* int16_t  */
      U32,  /* uint32_t */
      S32,  /* int32_t  */
      F32,  /* float    */
      F64,  /* double   */
      NbDataTypes
   } Type;
};

#endif /* __SIMAGE__H */
