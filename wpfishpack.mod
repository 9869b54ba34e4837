﻿!mod$ v1 sum:a3dfadef75a0caaf
module wpfishpack
integer(4),parameter::wp=8_4
end
